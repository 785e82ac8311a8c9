// Host side of the MI355X-native propagation path: the context behind the C ABI
// of include/tdz.h, the matrix-free operator apply, the preconditioners, FGMRES
// and the Crank-Nicolson driver.
//
// Reference behaviour restated here (C++ over HIP kernels, no PETSc):
//   TSTHETA theta=1/2 CN step, midpoint field     src/propagator.cpp:94-101, src/monitor.cpp:150-162, :367-382
//   FGMRES, restart 30 / 50 above 1e5 rows, right PC, nonzero guess
//                                                 src/propagator.cpp:131-182
//   PC refresh when its > 2*base + 5              include/tdsez_propagator.hpp:17-31, src/propagator.cpp:22-33
//   GMG hierarchy, Boehm P, R = P^T, Galerkin     src/multigrid.cpp:205-327
//   H = K + s V + CAP, field terms s E_a D_a      include/tdsez_parser.hpp:92-105, :190-194; src/manager.cpp:1011-1027
//   monitors: norm, <H>, dipoles                  src/manager.cpp:584-672, src/monitor.cpp:1145-1194
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tdz.h"
#include "tdz_internal.hpp"

using cplx = std::complex<double>;
using tdz::cpx;
using tdz::dcomplex;
using tdz::fcomplex;

namespace {

thread_local std::string g_err;
thread_local std::string g_warn;   // the last warning of the calling thread

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) throw Error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + #x); \
    } while (0)

inline dcomplex dc(cplx z) { return {z.real(), z.imag()}; }

template <typename T>
T *dalloc(size_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T *>(p);
}

// Sparse 1D matrix in CSR on device (values in both precisions).
struct CSR {
    int rows = 0;
    int *rp = nullptr, *col = nullptr;
    double *val = nullptr;
    float *valf = nullptr;
};

// Device data of one level in one precision.
template <typename R>
struct Prec {
    R *B0[3] = {}, *X[3] = {}, *W[3] = {};   // basis tables, points, weights
    R *M[3] = {};                            // real mass bands
    cpx<R> *B[3] = {};                       // complex combined bands (current shift)
    cpx<R> *Bmass[3] = {};                   // bands of M alone (M on x, zero on y and z)
    R *V[3] = {}, *lam[3] = {};              // separable preconditioner
    R *Vf[3][2] = {};                        // complex64: V and V^T in MFMA fragment order (or null)
    unsigned short *Vh[3][2] = {};           // complex64: V and V^T split into FP16 halves (or null)
    float Vh_inv[3][2] = {};                 // the power of two their scaling undoes
    R *Mw = nullptr;                         // layered z axis: 1/m-weighted mass band
    cpx<R> *lwork = nullptr;                 // layered z axis: work of the z-line solves
    cpx<R> *r = nullptr, *z = nullptr, *t = nullptr;   // multigrid work vectors
    // Jacobi-Chebyshev smoother: 1 / diag(A) for the shift of version dinv_ver, the
    // recurrence's direction, residual, operator product and sum, the lumped finest-level
    // potential s V(Greville)
    cpx<R> *dinv = nullptr, *cd = nullptr, *cres = nullptr, *cad = nullptr, *ce = nullptr;
    R *Vl = nullptr;
    unsigned dinv_ver = ~0u;
    bool ready = false;
};

struct Level {
    tdz::Axis1D ax[3];
    int n[3] = {1, 1, 1};
    long N = 1;
    int *dOff[3] = {};
    unsigned char *dWall[3] = {};
    // host bands of this level (Galerkin on coarse levels)
    std::vector<double> M[3], S[3], Xb[3], Wc[3], Vs[3];
    // layered z axis: bands of 1/m, (1/m) B'B' and the layer potential (Galerkin on coarse levels)
    bool layered = false;
    std::vector<double> Mw, Sw, Vl;
    // host eigen-decomposition of the separable preconditioner
    std::vector<double> hV[3], hLam[3];
    bool fd_ready = false;
    bool stale = false;     // coarse level whose bands wait for the next V-cycle
    // host staging of the band uploads (kept so that no copy waits on its source)
    std::vector<dcomplex> hB[3];
    std::vector<fcomplex> hBf[3], hBc[3];
    // device copies of the stiffness, CAP and position bands (bands updated on the device)
    double *dS[3] = {}, *dW[3] = {}, *dX[3] = {};
    Prec<double> d;
    Prec<float> f;
    // transfer to the next coarser level: P (this x coarse) and R = P^T
    CSR P[3], R[3];
    // dim 3 with element e starting at function e on every axis (tiled kernel applies)
    bool simple = false;    // 3D, simple knots on every axis: the tiled 3D potential kernel
    bool simple2 = false;   // 2D, simple knots: the tiled 2D potential kernel
    int pot_variant = -1;   // tile shape of the tiled kernel picked by the autotuner (-1: generic)
    double cheb_max = 0;    // Chebyshev smoother: upper eigenvalue bound of D^-1 M (0: not set up)
    // coarse levels: the exact Galerkin potential P^T (s V) P as an offset-major stencil
    double *Sv = nullptr;
    float *Svf = nullptr;
    long Sv_len = 0;
};

template <typename R> Prec<R> &prec(Level &L);
template <> Prec<double> &prec<double>(Level &L) { return L.d; }
template <> Prec<float> &prec<float>(Level &L) { return L.f; }

template <typename R> const R *csr_val(const CSR &s);
template <> const double *csr_val<double>(const CSR &s) { return s.val; }
template <> const float *csr_val<float>(const CSR &s) { return s.valf; }

// fine-level scratch of one precision
template <typename R>
struct Scratch {
    cpx<R> *tmp[4] = {};
    cpx<R> *w1 = nullptr, *w2 = nullptr;
    cpx<R> *u1 = nullptr, *u2 = nullptr;
    cpx<R> *y0 = nullptr;   // corrected preconditioners: y = S^-1 r across the sweeps
};

}  // namespace

namespace {
// The learned part of the preconditioner: a small 3D CNN that reproduces the part of
// the potential product V y the lumped form Vg .* M y misses (see tdz_learn.hip).
struct LearnState {
    bool ready = false;
    float *P = nullptr;                  // [N][C] positional fields
    tdz::CnnFieldGen fgen;               // the same fields' closed forms (inference)
    std::vector<double> fscale;          // [C] maximum of each field before scaling to 1
    float *W1 = nullptr, *W2 = nullptr;  // device weights [27][CI][C], [27][C][C]
    float *gr[3] = {};                   // Greville coordinates per axis (float)
    float *Xr[3] = {};                   // position band per axis (float)
    fcomplex *feat = nullptr;            // [CI][N] network input: y and its first-moment features
    tdz::CnnScales sin{{1.f, 1.f, 1.f, 1.f}};   // input scales (unit mean square over the samples)
    double sout = 1.0;                   // target scale: the network outputs sout * t
    std::vector<float> hW1, hW2;         // host weights
    unsigned wver = 0, packed_ver = ~0u;  // version of the device weights; the one packed for inference
    unsigned short *mb = nullptr;        // inference work [2N][C] bf16
    float *out = nullptr;                // [2N]
    // training
    float *m = nullptr, *dov = nullptr, *dh = nullptr, *partial = nullptr, *dW1 = nullptr, *dW2 = nullptr;
    double *lpart = nullptr;             // loss partial sums [kLossBlocks]
    std::vector<fcomplex *> ys, ts;      // harvested samples: network input [CI][N] and target t
    std::vector<double> tnorm2;          // |t|^2 per sample
    int max_samples = 0;
    bool harvesting = false;
    bool trained = false;
    int base_its = -1;    // iterations of the same preconditioner without the network
    int fallbacks = 0;
    bool fallen = false;  // the network regressed: the fallback preconditioner is in use
    int fallback_kind = TDZ_PC_MG;
    int random_count = 6;                // synthetic samples added to the harvested ones
    unsigned long long seed = 12345;
};

// the host weights to the device; a new version for the inference's packed operands
void upload_weights(LearnState &ls)
{
    HIPCHK(hipMemcpy(ls.W1, ls.hW1.data(), ls.hW1.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ls.W2, ls.hW2.data(), ls.hW2.size() * 4, hipMemcpyHostToDevice));
    ++ls.wver;
}
}  // namespace

struct tdz_ctx {
    int device = -1;            // the HIP device current at creation (its stream's device)
    tdz_problem prob{};
    tdz::PhysParams ph;
    hipStream_t st = nullptr;
    std::vector<std::unique_ptr<Level>> lv;
    cplx alpha = 1.0, g = 0.0;
    double sE[3] = {0, 0, 0};
    Scratch<double> sd;
    Scratch<float> sf;
    bool mixed_ready = false;   // single-precision copies allocated
    dcomplex *rhs = nullptr;
    // layered structure along z: bounds, effective masses, potential offsets
    bool layered = false;
    std::vector<double> lay_b, lay_m, lay_v;
    // captured CN step (batched scan): device-side FGMRES state, this run's field per
    // step, step counter, residual estimate per step, the field-dependent bands
    struct Captured {
        tdz::GmState *gm = nullptr;
        double *bnorm2 = nullptr, *sE = nullptr, *res = nullptr;
        int *step = nullptr;
        dcomplex *coef = nullptr, *T0[3] = {}, *bandR[3] = {}, *bandL[3] = {};
        double *Xb[3] = {}, *M0 = nullptr;
        int nsteps = 0, k = 0;
        double dt = 0;
        int pred = 0;                          // extrapolation order of the guess (0: the state)
        dcomplex *h1 = nullptr, *h2 = nullptr;   // the two states before the current one
    } cap;
    // CN predictor: the last states of a run with their field-free operator products and
    // dipole products, for the extrapolated initial guess and its residual
    struct Predictor {
        int order = 0;              // extrapolation order of the guess (0: the current state)
        int count = 0, head = 0;    // states held (at most 3), slot of the newest
        double dt = 0;
        unsigned axes = 0;          // field axes whose dipole products are kept
        static constexpr int kSlots = 5;   // order <= 4
        int ring = 1;                      // slots in use: order + 1
        int sync_every = 16, since_sync = 0;   // exact recomputation of the stored products
        double last_drift = 0;                 // their relative drift found at the last one
        dcomplex *psi[kSlots] = {}, *P[kSlots] = {}, *D[kSlots][3] = {};   // state; (M + g H_sEh) state; D_a state
        double sEh[kSlots][3] = {};                                        // the field of each slot's P
        dcomplex *Mx[kSlots] = {};                           // M state
        dcomplex *Xband[3] = {}, *Zband[3] = {};   // complex position bands and zero bands (dipole applies)
        double *Xz = nullptr;                      // the real z position band (fused mass + z dipole)
    } pred;
    dcomplex *zerov = nullptr;   // a zero vector (wall rows of the field terms)
    int restart = 30;
    dcomplex *Vb = nullptr, *Zb = nullptr;
    dcomplex *dpart = nullptr, *dh = nullptr, *dh2 = nullptr;
    double *dpartr = nullptr, *dnorm = nullptr;   // dnorm: 2 slots
    // pinned staging of the solver's small transfers (host memory that outlives the copies)
    dcomplex *hstage = nullptr;
    int hstage_len = 0;
    // the caller queued dnorm[0] = ||V_0||^2 and dnorm[1] = ||b||^2 for the next solve
    bool norms_queued = false;
    int pc_kind = TDZ_PC_NONE;
    bool pc_ready = false;
    std::vector<void *> allocs;     // everything to free
    std::vector<double *> user_vecs;
    int refresh_base = -1, rebuilds = 0;
    int stage_base[2] = {-1, -1};   // refresh trackers of the two Pade stages
    bool field_warned = false, cap_warned = false;   // the multigrid warnings were given
    double field_checked[8] = {};                      // the (pulse, tau, dt) the field check last passed
    int refresh_cap = -1;
    int smoother = 0;           // multigrid smoother: 0 separable-solve Richardson, 1 Jacobi-Chebyshev
    int fd_transform = 1;       // complex64 separable solves: 0 FP32 MFMA, 1 split FP16 MFMA
    int corr_form = 1;          // correction sweeps: 0 solve for the whole iterate, 1 for the correction alone
    bool pc_cform = false;      // this adaptive solve takes the correction form (tol / r0 below 1e-6)
    bool fd64_mfma = true;      // FP64 separable solves: FP64 MFMA transforms (else VALU, LDS-tiled)
    int cheb_degree = 2;
    // coarsest multigrid level: 0 separable-solve sweeps, 1 Krylov solve to coarse_rtol
    int coarse_kind = 0, coarse_maxit = 64;
    double coarse_rtol = 1e-12;
    struct CoarseKrylov {       // workspace of the coarse Krylov solve (allocated on first use)
        dcomplex *V = nullptr, *Z = nullptr, *dpart = nullptr, *dh = nullptr, *y = nullptr, *b = nullptr, *x = nullptr;
        double *dpartr = nullptr, *bnorm2 = nullptr, *res = nullptr;
        int *step = nullptr;
        tdz::GmState *gm = nullptr;
        int last_its = 0;
        double last_res = 0;
    } ck;
    unsigned shift_ver = 0;     // incremented by every set_shift
    int corrections = 0;        // correction sweeps of the corrected and learned preconditioners (0: adaptive)
    // adaptive sweeps: the solve's initial relative residual and tolerance (set by FGMRES
    // before its first preconditioner apply), the sweeps chosen for it, and the last
    // observed mean contraction per iteration of each choice
    double pc_r0rel = -1, pc_rtol = 0;
    bool pc_choose = false;
    int pc_sweeps = 1;
    bool pc_net = true;        // the learned preconditioner runs its network (in the last sweep)
    double rho_hat[2][5] = {{-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}};   // [network][sweeps]
    long sweep_count[2][4] = {};   // corrected solves per [network][sweeps] choice since pc_setup
    double its_hat[2][4] = {};     // moving average of the Krylov iterations of each choice
    double its_lt[2][4] = {};      // and of log10(tol / r0) of the solves it averages
    long pc_solves = 0;         // adaptive solves since pc_setup (the exploration schedule)
    long pc_next_drop = 12;     // the next solve that tries one sweep fewer
    int pc_drop_iv = 32;        // solves between such tries (doubled after each that costs iterations)
    bool pc_exploring = false;  // this solve tries one sweep fewer than the model's choice
    long pc_next_flip = 4;      // the next solve that tries the other network choice
    int pc_flip_iv = 32;        // solves between such tries (doubled after each that gained nothing)
    bool pc_flipping = false;   // this solve tries the other network choice
    int pc_flip_from_k = 0;     // the model's choice that solve replaced: sweeps and network
    bool pc_flip_from_net = false;
    bool cap_on = true;         // the CAP is part of H (off during the eigensolve, as src/propagator.cpp:51-69)
    // Pade(2,2) stage guesses: per slot (2 x stage) the last increments y - x
    dcomplex *hist = nullptr;   // [4 slots][3] vectors
    int hist_count[4] = {0, 0, 0, 0};
    int guess_order = 3, hist_order = 0;   // -pade_guess_order (1..5); order of the allocated history
    bool guess_extrap = true;              // -pade_extrapolate_guess
    dcomplex *xin = nullptr;    // stage input copy
    bool use_tiled = true;   // tiled 3D potential kernel where it applies
    // first-order corrected separable preconditioner: the non-separable part of s V at the
    // Greville points (lumped product) and the bands of g (CAP + sE.D), fine level, complex64
    float *dVg = nullptr;
    fcomplex *dBcorr[3] = {};
    dcomplex *pot_scr = nullptr;   // scratch of the tiled 3D potential kernel (finest level)
    fcomplex *pot_scrf = nullptr;
    LearnState learn;
    // kernel timing (measurement hook)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;
};

namespace {

template <typename R> Scratch<R> &scratch(tdz_ctx *c);
template <> Scratch<double> &scratch<double>(tdz_ctx *c) { return c->sd; }
template <> Scratch<float> &scratch<float>(tdz_ctx *c) { return c->sf; }

template <typename T>
T *track(tdz_ctx *c, T *p)
{
    c->allocs.push_back(p);
    return p;
}

// free a tracked allocation before the context ends
void untrack(tdz_ctx *c, void *p)
{
    auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
    if (it == c->allocs.end()) return;
    c->allocs.erase(it);
    HIPCHK(hipFree(p));
}

template <typename T>
T *upload(tdz_ctx *c, const std::vector<T> &h)
{
    T *d = track(c, dalloc<T>(h.size()));
    if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

std::vector<float> to_float(const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); }

// B = P^T A P for a band A (fine level l) and sparse P (fine x coarse), returned as a coarse band.
std::vector<double> galerkin_band(const std::vector<double> &A, int nf, int p,
                                  const std::vector<std::vector<std::pair<int, double>>> &P, int nc)
{
    const int bw = 2 * p + 1;
    // AP (nf x nc) dense, then P^T (AP)
    std::vector<double> AP((size_t)nf * nc, 0.0);
    for (int i = 0; i < nf; ++i)
        for (int j = std::max(0, i - p); j <= std::min(nf - 1, i + p); ++j) {
            const double a = A[(size_t)i * bw + (j - i + p)];
            if (a == 0.0) continue;
            for (auto &e : P[j]) AP[(size_t)i * nc + e.first] += a * e.second;
        }
    std::vector<double> C((size_t)nc * bw, 0.0);
    for (int i = 0; i < nf; ++i)
        for (auto &e : P[i])
            for (int J = std::max(0, e.first - p); J <= std::min(nc - 1, e.first + p); ++J)
                C[(size_t)e.first * bw + (J - e.first + p)] += e.second * AP[(size_t)i * nc + J];
    return C;
}

CSR make_csr(tdz_ctx *c, const std::vector<std::vector<std::pair<int, double>>> &rows)
{
    std::vector<int> rp(1, 0), col;
    std::vector<double> val;
    for (auto &r : rows) {
        for (auto &e : r) { col.push_back(e.first); val.push_back(e.second); }
        rp.push_back((int)col.size());
    }
    CSR s;
    s.rows = (int)rows.size();
    s.rp = upload(c, rp);
    s.col = upload(c, col);
    s.val = upload(c, val);
    s.valf = upload(c, to_float(val));
    return s;
}

// M alone as one Kronecker term: M on x (zero y and z bands), or on z for a layered
// structure, whose x and y terms carry the 1/m-weighted z band
std::vector<dcomplex> mass_only_band(const Level &L, int d)
{
    std::vector<dcomplex> bm(L.M[d].size(), dcomplex{0.0, 0.0});
    if (d == (L.layered ? 2 : 0))
        for (size_t k = 0; k < bm.size(); ++k) bm[k] = {L.M[d][k], 0.0};
    return bm;
}

void upload_level_tables(tdz_ctx *c, Level &L)
{
    for (int d = 0; d < 3; ++d) {
        const tdz::Axis1D &A = L.ax[d];
        L.d.B0[d] = upload(c, A.B0);
        L.d.X[d] = upload(c, A.X);
        L.d.W[d] = upload(c, A.W);
        L.dOff[d] = upload(c, A.off);
        L.d.M[d] = upload(c, L.M[d]);
        std::vector<unsigned char> w(A.isWall.begin(), A.isWall.end());
        L.dWall[d] = upload(c, w);
        L.d.B[d] = track(c, dalloc<dcomplex>(L.M[d].size()));
        L.d.Bmass[d] = upload(c, mass_only_band(L, d));
    }
    if (L.layered) L.d.Mw = upload(c, L.Mw);
    L.d.t = track(c, dalloc<dcomplex>(L.N));
    if (&L != c->lv[0].get()) {
        L.d.r = track(c, dalloc<dcomplex>(L.N));
        L.d.z = track(c, dalloc<dcomplex>(L.N));
    }
    L.d.ready = true;
}

// Single-precision copies of one level (tables, bands, work vectors), for the
// mixed-precision preconditioners. The separable preconditioner's eigenvectors
// are copied by setup_fd.
void upload_fd_float(tdz_ctx *c, Level &L, int d);

void upload_level_tables_f(tdz_ctx *c, Level &L)
{
    if (L.f.ready) return;
    for (int d = 0; d < 3; ++d) {
        const tdz::Axis1D &A = L.ax[d];
        L.f.B0[d] = upload(c, to_float(A.B0));
        L.f.X[d] = upload(c, to_float(A.X));
        L.f.W[d] = upload(c, to_float(A.W));
        L.f.M[d] = upload(c, to_float(L.M[d]));
        L.f.B[d] = track(c, dalloc<fcomplex>(L.M[d].size()));
        std::vector<dcomplex> bm = mass_only_band(L, d);
        std::vector<fcomplex> bf(bm.size());
        for (size_t k = 0; k < bm.size(); ++k) bf[k] = {(float)bm[k].re, (float)bm[k].im};
        L.f.Bmass[d] = upload(c, bf);
        if (L.fd_ready) upload_fd_float(c, L, d);
    }
    if (L.layered) {
        L.f.Mw = upload(c, to_float(L.Mw));
        if (L.fd_ready) L.f.lwork = track(c, dalloc<fcomplex>((size_t)L.N * (2 * L.ax[2].p + 1)));
    }
    L.f.t = track(c, dalloc<fcomplex>(L.N));
    L.f.r = track(c, dalloc<fcomplex>(L.N));
    L.f.z = track(c, dalloc<fcomplex>(L.N));
    L.f.ready = true;
}

// Complex bands of alpha M + g (K + CAP + sE.D) split over the three axes:
// B_x = alpha M_x + g T_x, B_y = g T_y, B_z = g T_z, T_d = S_d/(2m) - i W_d + sE_d X_d.
void update_bands(tdz_ctx *c, Level &L)
{
    const double inv2m = 0.5 / c->ph.mass;
    bool on_device = !L.layered;
    for (int d = 0; d < 3 && on_device; ++d) {
        const size_t sz = L.M[d].size();
        on_device = L.d.M[d] && L.d.B[d] && L.S[d].size() == sz && L.Wc[d].size() == sz && L.Xb[d].size() == sz;
    }
    if (on_device) {
        // on the device: no host staging, no copies, no wait for the stream
        tdz::BandUpdate bu;
        for (int d = 0; d < 3; ++d) {
            const size_t sz = L.M[d].size();
            if (!L.dS[d]) {
                L.dS[d] = upload(c, L.S[d]);
                L.dW[d] = upload(c, L.Wc[d]);
                L.dX[d] = upload(c, L.Xb[d]);
            }
            bu.M[d] = L.d.M[d];
            bu.S[d] = L.dS[d];
            bu.W[d] = L.dW[d];
            bu.X[d] = L.dX[d];
            bu.B[d] = L.d.B[d];
            bu.Bc[d] = (c->dBcorr[d] && &L == c->lv[0].get()) ? c->dBcorr[d] : nullptr;
            bu.Bf[d] = L.f.ready ? L.f.B[d] : nullptr;
            bu.sz[d] = (int)sz;
            bu.active[d] = d < c->prob.dim;
            bu.sE[d] = c->sE[d];
        }
        bu.alpha = dc(c->alpha);
        bu.g = dc(c->g);
        bu.inv2m = inv2m;
        bu.cap = c->cap_on ? 1 : 0;
        tdz::launch_band_update(c->st, bu);
        return;
    }
    for (int d = 0; d < 3; ++d) {
        const size_t sz = L.M[d].size();
        std::vector<dcomplex> &B = L.hB[d];
        B.resize(sz);
        const bool active = d < c->prob.dim;
        for (size_t k = 0; k < sz; ++k) {
            cplx b;
            if (L.layered) {
                // x, y: g S/2 (their z factor is the 1/m band, Mz2); z: alpha M + g (Sw/2 + s V_l + sE_z X)
                if (d < 2) b = c->g * (0.5 * L.S[d][k]);
                else b = c->alpha * L.M[2][k] + c->g * (0.5 * L.Sw[k] + c->ph.s * L.Vl[k] + c->sE[2] * L.Xb[2][k]);
            } else {
                cplx T = 0.0;
                if (active)
                    T = inv2m * L.S[d][k] + (c->cap_on ? cplx(0, -1) * L.Wc[d][k] : 0.0) + c->sE[d] * L.Xb[d][k];
                b = c->g * T;
                if (d == 0) b += c->alpha * L.M[d][k];
            }
            B[k] = dc(b);
        }
        HIPCHK(hipMemcpyAsync(L.d.B[d], B.data(), sz * sizeof(dcomplex), hipMemcpyHostToDevice, c->st));
        if (c->dBcorr[d] && &L == c->lv[0].get()) {
            std::vector<fcomplex> &Bc = L.hBc[d];
            Bc.assign(sz, fcomplex{0.f, 0.f});
            if (active)
                for (size_t k = 0; k < sz; ++k) {
                    const cplx b = c->g * ((c->cap_on ? cplx(0, -1) * L.Wc[d][k] : 0.0) + c->sE[d] * L.Xb[d][k]);
                    Bc[k] = {(float)b.real(), (float)b.imag()};
                }
            HIPCHK(hipMemcpyAsync(c->dBcorr[d], Bc.data(), sz * sizeof(fcomplex), hipMemcpyHostToDevice, c->st));
        }
        if (L.f.ready) {
            std::vector<fcomplex> &Bf = L.hBf[d];
            Bf.resize(sz);
            for (size_t k = 0; k < sz; ++k) Bf[k] = {(float)B[k].re, (float)B[k].im};
            HIPCHK(hipMemcpyAsync(L.f.B[d], Bf.data(), sz * sizeof(fcomplex), hipMemcpyHostToDevice, c->st));
        }
    }
    HIPCHK(hipStreamSynchronize(c->st));   // the staging vectors are reused by the next update
}

// Coarse levels are only read by the V-cycle: while the preconditioner does not use
// them their bands are left stale and brought up to date by refresh_coarse.
bool uses_coarse(const tdz_ctx *c)
{
    return c->pc_kind == TDZ_PC_MG || c->pc_kind == TDZ_PC_MG_MIXED ||
           (c->pc_kind == TDZ_PC_LEARNED && c->learn.fallen);
}

void set_shift_bands(tdz_ctx *c)
{
    const bool coarse = uses_coarse(c);
    for (size_t l = 0; l < c->lv.size(); ++l) {
        Level &L = *c->lv[l];
        if (l > 0 && !coarse) {
            L.stale = true;
            continue;
        }
        update_bands(c, L);
        L.stale = false;
    }
}

void refresh_coarse(tdz_ctx *c)
{
    for (size_t l = 1; l < c->lv.size(); ++l)
        if (c->lv[l]->stale) {
            update_bands(c, *c->lv[l]);
            c->lv[l]->stale = false;
        }
}

template <typename R>
tdz::AxisDevT<R> axis_dev(Level &L, int d)
{
    const tdz::Axis1D &A = L.ax[d];
    Prec<R> &P = prec<R>(L);
    return {A.E, A.p, A.Q, A.n, P.B0[d], P.X[d], P.W[d], L.dOff[d]};
}

template <typename R>
tdz::SepBandsT<R> sep_bands(Level &L)
{
    tdz::SepBandsT<R> sb;
    Prec<R> &P = prec<R>(L);
    for (int d = 0; d < 3; ++d) {
        sb.n[d] = L.n[d];
        sb.p[d] = L.ax[d].p;
        sb.M[d] = P.M[d];
        sb.B[d] = P.B[d];
    }
    sb.Mz2 = L.layered ? P.Mw : nullptr;
    return sb;
}

template <typename R>
inline cpx<R> cx(cplx z) { return {(R)z.real(), (R)z.imag()}; }

// Scratch of the tiled 3D potential kernel (the finest level only; allocated at first use).
template <typename R> cpx<R> *&pot_scratch_slot(tdz_ctx *c);
template <> dcomplex *&pot_scratch_slot<double>(tdz_ctx *c) { return c->pot_scr; }
template <> fcomplex *&pot_scratch_slot<float>(tdz_ctx *c) { return c->pot_scrf; }
template <typename R>
cpx<R> *pot_scratch(tdz_ctx *c, const Level &L)
{
    if (&L != c->lv[0].get() || !L.simple) return nullptr;
    cpx<R> *&s = pot_scratch_slot<R>(c);
    if (!s) s = track(c, dalloc<cpx<R>>(tdz::pot3d_scratch_len(L.N)));
    return s;
}

// y += g (s V) x on level L: the tiled 3D kernel where it applies, else the generic one.
template <typename R>
void pot_apply(tdz_ctx *c, Level &L, cplx g, const cpx<R> *x, cpx<R> *y)
{
    tdz::AxisDevT<R> ax[3] = {axis_dev<R>(L, 0), axis_dev<R>(L, 1), axis_dev<R>(L, 2)};
    tdz::PotParams pp{c->ph.pot_kind, c->ph.s, c->ph.pot[0], c->ph.pot[1]};
    const bool timed = std::is_same<R, double>::value && c->timing && &L == c->lv[0].get();
    if (timed) {
        if (c->ev_used == c->ev.size()) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            c->ev.push_back({a, b});
        }
        HIPCHK(hipEventRecord(c->ev[c->ev_used].first, c->st));
    }
    const bool tiled =
        c->use_tiled &&
        ((L.simple && tdz::launch_pot_apply_3d<R>(c->st, ax, pp, (R)g.real(), (R)g.imag(), x, y, L.pot_variant,
                                                  pot_scratch<R>(c, L))) ||
         (L.simple2 && tdz::launch_pot_apply_2d<R>(c->st, ax, pp, (R)g.real(), (R)g.imag(), x, y, L.pot_variant)));
    if (!tiled) tdz::launch_pot_apply_generic<R>(c->st, ax, pp, (R)g.real(), (R)g.imag(), x, y);
    if (timed) HIPCHK(hipEventRecord(c->ev[c->ev_used++].second, c->st));
}

void ensure_galerkin(tdz_ctx *c);

template <typename R> const R *stencil_of(const Level &L);
template <> const double *stencil_of<double>(const Level &L) { return L.Sv; }
template <> const float *stencil_of<float>(const Level &L) { return L.Svf; }

// y = A_l x with the current shift: separable bands (Galerkin products per axis on coarse
// levels) plus the potential, matrix-free on the finest level and the exact Galerkin
// stencil P^T (s V) P on the coarse ones.
template <typename R>
void level_apply(tdz_ctx *c, Level &L, const cpx<R> *x, cpx<R> *y)
{
    tdz::launch_sep_apply<R>(c->st, sep_bands<R>(L), x, y, scratch<R>(c).tmp);
    if (c->ph.pot_kind && c->g != 0.0) {
        if (&L == c->lv[0].get()) {
            pot_apply<R>(c, L, c->g, x, y);
        } else {
            if (!stencil_of<R>(L)) ensure_galerkin(c);
            const int hp[3] = {L.ax[0].p, L.ax[1].p, L.ax[2].p};
            tdz::launch_stencil_apply<R>(c->st, L.n, hp, stencil_of<R>(L), (R)c->g.real(), (R)c->g.imag(), x, y);
        }
    }
    if (c->prob.dirichlet) {
        const unsigned char *w[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
        tdz::launch_wall_fix<R>(c->st, L.n, w, x, y);
    }
}

// Autotuner: time every tile-shape variant of the tiled potential kernel on this
// level's grid (one warm-up and two timed launches each, HIP events on the
// context stream) and keep the fastest. TDZ_POT_VARIANT=k forces variant k.
void autotune_pot(tdz_ctx *c, Level &L)
{
    L.pot_variant = -1;
    if (!(L.simple || L.simple2) || !c->ph.pot_kind) return;
    const int nv = L.simple ? tdz::pot3d_variants(L.ax[0].p, L.ax[0].Q) : tdz::pot2d_variants(L.ax[0].p, L.ax[0].Q);
    if (nv == 0) return;
    if (const char *f = std::getenv("TDZ_POT_VARIANT")) {
        const int k = std::atoi(f);
        L.pot_variant = (k >= 0 && k < nv) ? k : -1;
        return;
    }
    if (nv == 1) { L.pot_variant = 0; return; }
    dcomplex *x = dalloc<dcomplex>(L.N), *y = dalloc<dcomplex>(L.N);
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipMemsetAsync(x, 0, L.N * sizeof(dcomplex), c->st));
    HIPCHK(hipMemsetAsync(y, 0, L.N * sizeof(dcomplex), c->st));
    tdz::AxisDev ax[3] = {axis_dev<double>(L, 0), axis_dev<double>(L, 1), axis_dev<double>(L, 2)};
    tdz::PotParams pp{c->ph.pot_kind, c->ph.s, c->ph.pot[0], c->ph.pot[1]};
    float best = 1e30f;
    auto launch = [&](int v) {
        if (L.simple) tdz::launch_pot_apply_3d<double>(c->st, ax, pp, 1.0, 0.0, x, y, v, pot_scratch<double>(c, L));
        else tdz::launch_pot_apply_2d<double>(c->st, ax, pp, 1.0, 0.0, x, y, v);
    };
    // three rounds over the variants (each variant's best round counts), so that clock ramps
    // and one-off delays do not decide between variants a few percent apart
    std::vector<float> vbest(nv, 1e30f);
    for (int round = 0; round < 3; ++round)
        for (int v = 0; v < nv; ++v) {
            launch(v);
            HIPCHK(hipEventRecord(a, c->st));
            for (int r = 0; r < 2; ++r) launch(v);
            HIPCHK(hipEventRecord(b, c->st));
            HIPCHK(hipEventSynchronize(b));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, a, b));
            vbest[v] = std::min(vbest[v], ms);
        }
    for (int v = 0; v < nv; ++v)
        if (vbest[v] < best) { best = vbest[v]; L.pot_variant = v; }
    HIPCHK(hipEventDestroy(a));
    HIPCHK(hipEventDestroy(b));
    HIPCHK(hipFree(x));
    HIPCHK(hipFree(y));
}

// complex64 copies of the axis-d eigen-decomposition, with the fragment-ordered V and
// V^T of the register-resident MFMA transform (axes it takes)
void upload_fd_float(tdz_ctx *c, Level &L, int d)
{
    L.f.V[d] = upload(c, to_float(L.hV[d]));
    L.f.lam[d] = upload(c, to_float(L.hLam[d]));
    const int n = L.n[d];
    for (int tr = 0; tr < 2; ++tr) {
        std::vector<float> fr;
        if (L.hV[d].size() == (size_t)n * n) fr = tdz::axis_fragments(n, L.hV[d], tr == 1);
        L.f.Vf[d][tr] = fr.empty() ? nullptr : upload(c, fr);
        std::vector<unsigned short> fh;
        if (L.hV[d].size() == (size_t)n * n) fh = tdz::axis_fragments_split16(n, L.hV[d], tr == 1, L.f.Vh_inv[d][tr]);
        L.f.Vh[d][tr] = fh.empty() ? nullptr : upload(c, fh);
    }
}

void setup_fd(tdz_ctx *c, Level &L)
{
    if (L.fd_ready) return;
    const double inv2m = 0.5 / c->ph.mass;
    for (int d = 0; d < 3; ++d) {
        std::vector<double> V, lam;
        const tdz::Axis1D &A = L.ax[d];
        if (L.layered && d == 2) {   // z: solved along lines, no eigenbasis
            V = {1.0};
            lam = {0.0};
        } else if (d < c->prob.dim) {
            std::vector<double> T(L.S[d].size());
            for (size_t k = 0; k < T.size(); ++k)
                T[k] = L.layered ? 0.5 * L.S[d][k] : inv2m * L.S[d][k] + L.Vs[d][k];
            tdz::Axis1D Ab = A;
            Ab.M = L.M[d];
            tdz::gen_eig_bands(Ab, T, V, lam);
        } else {
            V = {1.0};
            lam = {0.0};
        }
        L.hV[d] = V;
        L.hLam[d] = lam;
        L.d.V[d] = upload(c, V);
        L.d.lam[d] = upload(c, lam);
        for (int tr = 0; tr < 2; ++tr) {   // the FP64 MFMA transform's fragments (axes it takes)
            std::vector<double> f;
            if (V.size() == (size_t)L.n[d] * L.n[d]) f = tdz::axis_fragments_f64(L.n[d], V, tr == 1);
            L.d.Vf[d][tr] = f.empty() ? nullptr : upload(c, f);
        }
        if (L.f.ready) upload_fd_float(c, L, d);
    }
    if (L.layered) {
        L.d.lwork = track(c, dalloc<dcomplex>((size_t)L.N * (2 * L.ax[2].p + 1)));
        if (L.f.ready) L.f.lwork = track(c, dalloc<fcomplex>((size_t)L.N * (2 * L.ax[2].p + 1)));
    }
    L.fd_ready = true;
}

// z = (alpha M + g (K + V_sep))^-1 r by fast diagonalisation (walls: identity);
// V_sep is the part of s V that is a sum of 1D terms. s1, s2: scratch.
// Layered structure: x and y by their eigenbases, z by a banded solve per (x, y) line of
// the whole separable operator there (the current z band, i.e. with the field along z,
// plus g (lambda_x + lambda_y) times the 1/m band): the exact inverse without a
// non-separable potential.
template <typename R>
void fd_apply(tdz_ctx *c, Level &L, const cpx<R> *r, cpx<R> *z, cpx<R> *s1, cpx<R> *s2, int kind = -1)
{
    Prec<R> &P = prec<R>(L);
    // the register-resident MFMA transforms' fragments (FP64: unless the VALU form is chosen)
    const bool frag = !std::is_same<R, double>::value || c->fd64_mfma;
    R *Vf[3][2];
    for (int d = 0; d < 3; ++d)
        for (int tr = 0; tr < 2; ++tr) Vf[d][tr] = frag ? P.Vf[d][tr] : nullptr;
    if (L.layered) {
        tdz::launch_dense_axis<R>(c->st, 0, L.n, P.V[0], true, r, s1, Vf[0][1]);
        tdz::launch_dense_axis<R>(c->st, 1, L.n, P.V[1], true, s1, s2, Vf[1][1]);
        tdz::launch_zline_solve<R>(c->st, L.n, L.ax[2].p, P.lam[0], P.lam[1], P.B[2], P.Mw, cx<R>(c->g), s2, s1,
                                   P.lwork);
        tdz::launch_dense_axis<R>(c->st, 0, L.n, P.V[0], false, s1, s2, Vf[0][0]);
        tdz::launch_dense_axis<R>(c->st, 1, L.n, P.V[1], false, s2, z, Vf[1][0]);
        return;
    }
    const R *lam[3] = {P.lam[0], P.lam[1], P.lam[2]};
    const unsigned char *w[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
    if constexpr (std::is_same<R, float>::value) {
        // split-FP16 transforms on every axis that takes them (the z pass with the diagonal solve)
        const int k16 = kind >= 0 ? kind : c->fd_transform;   // complex64 transform kind
        if (k16 > 0 && P.Vh[0][0] && P.Vh[1][0] && P.Vh[2][0] && tdz::split16_fits(L.n)) {
            tdz::launch_dense_axis_split16(c->st, 0, L.n, P.Vh[0][1], P.Vh_inv[0][1], r, s1, k16);
            tdz::launch_dense_axis_split16(c->st, 1, L.n, P.Vh[1][1], P.Vh_inv[1][1], s1, s2, k16);
            tdz::launch_dense_axis_diag_split16(c->st, L.n, P.Vh[2][1], P.Vh_inv[2][1], lam, w, cx<R>(c->alpha),
                                                cx<R>(c->g), s2, s1, k16);
            tdz::launch_dense_axis_split16(c->st, 0, L.n, P.Vh[0][0], P.Vh_inv[0][0], s1, s2, k16);
            tdz::launch_dense_axis_split16(c->st, 1, L.n, P.Vh[1][0], P.Vh_inv[1][0], s2, s1, k16);
            tdz::launch_dense_axis_split16(c->st, 2, L.n, P.Vh[2][0], P.Vh_inv[2][0], s1, z, k16);
            if (c->prob.dirichlet) tdz::launch_wall_fix<R>(c->st, L.n, w, r, z);
            return;
        }
    }
    tdz::launch_dense_axis<R>(c->st, 0, L.n, P.V[0], true, r, s1, Vf[0][1]);
    tdz::launch_dense_axis<R>(c->st, 1, L.n, P.V[1], true, s1, s2, Vf[1][1]);
    cpx<R> *d = s2, *e = s1;   // the diagonal solve's output and the other scratch vector
    bool fused = false;
    if constexpr (std::is_same<R, float>::value) {   // z pass and diagonal solve in one kernel
        fused = tdz::launch_dense_axis_diag(c->st, L.n, Vf[2][1], lam, w, cx<R>(c->alpha), cx<R>(c->g), s2, s1);
        if (fused) d = s1, e = s2;
    }
    if (!fused) {
        tdz::launch_dense_axis<R>(c->st, 2, L.n, P.V[2], true, s2, s1, Vf[2][1]);
        tdz::launch_diag_scale<R>(c->st, L.n, lam, w, cx<R>(c->alpha), cx<R>(c->g), s1, s2);
    }
    tdz::launch_dense_axis<R>(c->st, 0, L.n, P.V[0], false, d, e, Vf[0][0]);
    tdz::launch_dense_axis<R>(c->st, 1, L.n, P.V[1], false, e, d, Vf[1][0]);
    tdz::launch_dense_axis<R>(c->st, 2, L.n, P.V[2], false, d, z, Vf[2][0]);
    if (c->prob.dirichlet) tdz::launch_wall_fix<R>(c->st, L.n, w, r, z);
}

// z = restriction-to-coarse / prolongation along three axes (CSR passes)
template <typename R>
void transfer(tdz_ctx *c, const CSR ops[3], const int nin[3], const int nout[3], const cpx<R> *in, cpx<R> *out,
              cpx<R> *s1, cpx<R> *s2, bool accumulate)
{
    int a[3] = {nin[0], nin[1], nin[2]};
    tdz::launch_csr_axis<R>(c->st, 0, a, nout[0], ops[0].rp, ops[0].col, csr_val<R>(ops[0]), in, s1, false);
    a[0] = nout[0];
    tdz::launch_csr_axis<R>(c->st, 1, a, nout[1], ops[1].rp, ops[1].col, csr_val<R>(ops[1]), s1, s2, false);
    a[1] = nout[1];
    tdz::launch_csr_axis<R>(c->st, 2, a, nout[2], ops[2].rp, ops[2].col, csr_val<R>(ops[2]), s2, out, accumulate);
}

double dev_norm2(tdz_ctx *c, const dcomplex *x, long N);

// ---- Jacobi-Chebyshev smoother (the multigrid smoother of the mixed-precision
// configuration): a Chebyshev polynomial in D^-1 A, D = diag(A), over the interval
// [lambda_max / 30, lambda_max] of D^-1 M (M dominates A = M + i dt/2 H at CN steps; the
// bound comes from a power iteration on the level's mass operator at set-up).
template <typename R>
void jacobi_setup(tdz_ctx *c, Level &L)
{
    Prec<R> &P = prec<R>(L);
    const long N = L.N;
    if (!P.dinv) {
        P.dinv = track(c, dalloc<cpx<R>>(N));
        P.cd = track(c, dalloc<cpx<R>>(N));
        P.cres = track(c, dalloc<cpx<R>>(N));
        P.cad = track(c, dalloc<cpx<R>>(N));
        P.ce = track(c, dalloc<cpx<R>>(N));
    }
    const bool fine = &L == c->lv[0].get();
    if (fine && c->ph.pot_kind && !P.Vl) {   // s V at the Greville points (the lumped diagonal)
        std::vector<double> gr[3];
        for (int d = 0; d < 3; ++d)
            gr[d] = d < c->prob.dim ? tdz::greville(L.ax[d].t, L.ax[d].p) : std::vector<double>{0.0};
        std::vector<R> v(N);
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    const double r2 = gr[0][i] * gr[0][i] + gr[1][j] * gr[1][j] + gr[2][k] * gr[2][k];
                    const double V = c->ph.pot_kind == 1 ? 0.5 * c->ph.pot[0] * r2
                                                         : -c->ph.pot[0] / std::sqrt(r2 + c->ph.pot[1]);
                    v[i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k)] = (R)(c->ph.s * V);
                }
        P.Vl = upload(c, v);
    }
    if (P.dinv_ver == c->shift_ver) return;
    const R *M[3] = {P.M[0], P.M[1], P.M[2]};
    const cpx<R> *B[3] = {P.B[0], P.B[1], P.B[2]};
    const int p[3] = {L.ax[0].p, L.ax[1].p, L.ax[2].p};
    const unsigned char *wall[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
    const R *Vc = (!fine && c->ph.pot_kind) ? stencil_of<R>(L) : nullptr;
    tdz::launch_jacobi_diag<R>(c->st, L.n, p, M, B, Vc, fine ? P.Vl : nullptr, cx<R>(c->g), wall, P.dinv);
    P.dinv_ver = c->shift_ver;
}

// Upper eigenvalue bound of D_M^-1 M on level L (12 power iterations, FP64, 10% margin).
double cheb_bound(tdz_ctx *c, Level &L)
{
    const long N = L.N;
    jacobi_setup<double>(c, L);
    Prec<double> &P = L.d;
    const double *M[3] = {P.M[0], P.M[1], P.M[2]};
    const dcomplex *B[3] = {P.Bmass[0], P.Bmass[1], P.Bmass[2]};
    const int p[3] = {L.ax[0].p, L.ax[1].p, L.ax[2].p};
    const unsigned char *wall[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
    tdz::launch_jacobi_diag<double>(c->st, L.n, p, M, B, nullptr, nullptr, dcomplex{0, 0}, wall, P.cres);   // 1/diag M
    std::vector<dcomplex> h(N);
    for (long i = 0; i < N; ++i) h[i] = {std::sin(0.37 * (double)i + 0.1), std::cos(0.53 * (double)i)};
    HIPCHK(hipMemcpy(P.cd, h.data(), N * sizeof(dcomplex), hipMemcpyHostToDevice));
    tdz::SepBandsT<double> mb = sep_bands<double>(L);
    for (int d = 0; d < 3; ++d) mb.B[d] = P.Bmass[d];
    double lam = 0;
    for (int it = 0; it < 12; ++it) {
        tdz::launch_sep_apply<double>(c->st, mb, P.cd, P.cad, c->sd.tmp);
        tdz::launch_cheb_update<double>(c->st, N, 0.0, 1.0, P.cres, P.cad, P.ce, nullptr);   // D^-1 M v
        const double nv = dev_norm2(c, P.cd, N), nw = dev_norm2(c, P.ce, N);
        lam = std::sqrt(nw / nv);
        tdz::launch_axpby_t<double>(c->st, {1.0 / std::sqrt(nw), 0}, P.ce, {0, 0}, P.cd, N);
    }
    P.dinv_ver = ~0u;   // dinv's storage held 1/diag M above
    return 1.1 * lam;
}

// e = p(D^-1 A) D^-1 r, the degree-k Chebyshev approximation of A^-1 r (Saad, Iterative
// Methods for Sparse Linear Systems, Alg. 12.1, started from zero), k - 1 operator applies.
template <typename R>
void chebyshev(tdz_ctx *c, Level &L, const cpx<R> *r, cpx<R> *e)
{
    jacobi_setup<R>(c, L);
    Prec<R> &P = prec<R>(L);
    const long N = L.N;
    const double hi = L.cheb_max, lo = hi / 30.0;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    tdz::launch_cheb_update<R>(c->st, N, R(0), R(1.0 / theta), P.dinv, r, P.cd, nullptr);
    HIPCHK(hipMemcpyAsync(e, P.cd, N * sizeof(cpx<R>), hipMemcpyDeviceToDevice, c->st));
    const cpx<R> one{1, 0}, mone{-1, 0}, zero{0, 0};
    for (int k = 1; k < c->cheb_degree; ++k) {
        level_apply<R>(c, L, P.cd, P.cad);
        if (k == 1) tdz::launch_axpby_t<R>(c->st, one, r, zero, P.cres, N);   // res = r - A d
        tdz::launch_axpby_t<R>(c->st, mone, P.cad, one, P.cres, N);
        const double rn = 1.0 / (2.0 * sigma - rho);
        tdz::launch_cheb_update<R>(c->st, N, R(rn * rho), R(2.0 * rn / delta), P.dinv, P.cres, P.cd, e);
        rho = rn;
    }
}

// Richardson sweep z += S^-1 (r - A z) on level L (residual in the level's t)
template <typename R>
void smooth(tdz_ctx *c, Level &L, const cpx<R> *r, cpx<R> *z, cpx<R> *s1, cpx<R> *s2, cpx<R> *s3, bool cheb)
{
    cpx<R> *t = prec<R>(L).t;
    const cpx<R> one{1, 0}, mone{-1, 0};
    level_apply<R>(c, L, z, t);
    tdz::launch_axpby_t<R>(c->st, one, r, mone, t, L.N);  // t = r - A z
    if (cheb) {
        chebyshev<R>(c, L, t, prec<R>(L).ce);
        tdz::launch_axpby_t<R>(c->st, one, prec<R>(L).ce, one, z, L.N);
        return;
    }
    fd_apply<R>(c, L, t, s3, s1, s2);
    tdz::launch_axpby_t<R>(c->st, one, s3, one, z, L.N);   // z += S^-1 t
}

// Coarsest-level solve z = A_L^-1 r to the relative residual c->coarse_rtol: GMRES with the
// level's separable solve as right preconditioner, restarted every min(kGmMax, maxit)
// iterations, at most maxit iterations (rounded up to whole cycles), from z = 0. The Hessenberg least-squares problem stays on the device (gm_* kernels); the host
// reads the residual estimate once per cycle. This takes the place of the reference's exact
// coarse factorisation (KSPPREONLY + PCLU/MUMPS, src/multigrid.cpp:314-325): a dense or
// banded factor of the 36^3-function coarsest level of a 128^3 grid would hold 35 GB (dense)
// or 12 GB (band) and cost a pass over them per V-cycle, where a Krylov solve costs a few
// applies of the level's 729-point operator.
void coarse_krylov(tdz_ctx *c, Level &L, const dcomplex *r, dcomplex *z)
{
    constexpr int km = tdz::kGmMax;
    const int k = std::min(km, std::max(c->coarse_maxit, 1));   // iterations per cycle
    const long N = L.N;
    tdz_ctx::CoarseKrylov &K = c->ck;
    if (!K.V) {
        K.V = track(c, dalloc<dcomplex>((size_t)(km + 1) * N));
        K.Z = track(c, dalloc<dcomplex>((size_t)km * N));
        K.b = track(c, dalloc<dcomplex>(N));
        K.dpart = track(c, dalloc<dcomplex>((size_t)(km + 1) * tdz::kRedBlocks));
        K.dh = track(c, dalloc<dcomplex>(2 * (km + 1) + 1));
        K.y = track(c, dalloc<dcomplex>(km));
        K.dpartr = track(c, dalloc<double>(tdz::kRedBlocks));
        K.bnorm2 = track(c, dalloc<double>(1));
        K.res = track(c, dalloc<double>(1));
        K.step = track(c, dalloc<int>(1));
        K.gm = track(c, dalloc<tdz::GmState>(1));
        HIPCHK(hipMemsetAsync(K.step, 0, sizeof(int), c->st));
    }
    dcomplex *dh2 = K.dh + (km + 1);
    double *dnorm = reinterpret_cast<double *>(K.dh + 2 * (km + 1));
    Scratch<double> &S = scratch<double>(c);
    const dcomplex one{1, 0}, mone{-1, 0};
    HIPCHK(hipMemcpyAsync(K.b, r, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemsetAsync(z, 0, N * sizeof(dcomplex), c->st));
    tdz::launch_norm2(c->st, K.b, N, K.dpartr, K.bnorm2);
    int its = 0;
    double res = 1.0;
    for (int cycle = 0; its < std::max(c->coarse_maxit, 1); ++cycle) {
        dcomplex *V0 = K.V;
        if (cycle == 0) {
            HIPCHK(hipMemcpyAsync(V0, K.b, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
        } else {
            level_apply<double>(c, L, z, V0);
            tdz::launch_axpby_t<double>(c->st, one, K.b, mone, V0, N);   // V0 = b - A z
        }
        tdz::launch_norm2(c->st, V0, N, K.dpartr, dnorm);
        tdz::launch_gm_start(c->st, dnorm, K.bnorm2, K.gm);
        tdz::launch_scale_invsqrt(c->st, V0, dnorm, V0, N);
        for (int j = 0; j < k; ++j) {
            dcomplex *Vj = K.V + (long)j * N, *Zj = K.Z + (long)j * N, *w = K.V + (long)(j + 1) * N;
            fd_apply<double>(c, L, Vj, Zj, S.w1, S.w2);
            level_apply<double>(c, L, Zj, w);
            tdz::launch_multi_dot(c->st, K.V, N, j + 1, w, N, K.dpart, K.dh);
            tdz::launch_multi_axpy_norm(c->st, K.V, N, j + 1, K.dh, w, N, K.dpartr, dnorm);
            tdz::launch_multi_dot(c->st, K.V, N, j + 1, w, N, K.dpart, dh2);
            tdz::launch_multi_axpy_norm(c->st, K.V, N, j + 1, dh2, w, N, K.dpartr, dnorm);
            tdz::launch_gm_arnoldi(c->st, j, k, K.dh, dh2, dnorm, K.gm);
            tdz::launch_scale_invsqrt(c->st, w, dnorm, w, N);
        }
        tdz::launch_gm_solve(c->st, k, k, K.gm, K.y, K.res, K.step);
        tdz::launch_multi_axpy(c->st, K.Z, N, k, K.y, z, N);
        its += k;
        HIPCHK(hipMemcpyAsync(&res, K.res, sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (!(res > c->coarse_rtol)) break;
    }
    K.last_its = its;
    K.last_res = res;
}

// V-cycle z = B_l r on level l (z is overwritten), in precision R.
template <typename R>
void vcycle(tdz_ctx *c, int l, const cpx<R> *r, cpx<R> *z)
{
    Level &L = *c->lv[l];
    const int nu = std::max(1, c->prob.mg_smooth);
    // scratch: w1, w2 (fine-level sized) for the transforms; tmp[3] holds a smoother
    // correction (level_apply, the only other user of tmp, has finished by then)
    Scratch<R> &S = scratch<R>(c);
    cpx<R> *s1 = S.w1, *s2 = S.w2;
    const cpx<R> one{1, 0}, mone{-1, 0};
    const bool coarsest = l + 1 == (int)c->lv.size();
    const bool cheb = c->smoother == 1 && !coarsest;   // the coarsest level keeps its separable sweeps
    if (coarsest && l > 0 && c->coarse_kind == 1) {    // Krylov solve, in FP64 for either precision
        if constexpr (std::is_same<R, double>::value) {
            coarse_krylov(c, L, r, z);
        } else {
            tdz::launch_promote(c->st, r, L.d.z, L.N);
            coarse_krylov(c, L, L.d.z, L.d.r);
            tdz::launch_demote(c->st, L.d.r, z, L.N);
        }
        return;
    }
    if (cheb) chebyshev<R>(c, L, r, z);
    else fd_apply<R>(c, L, r, z, s1, s2);
    const int pre = coarsest ? nu + 1 : nu;
    for (int s = 1; s < pre; ++s) smooth<R>(c, L, r, z, s1, s2, S.tmp[3], cheb);
    if (coarsest) return;
    Level &Cl = *c->lv[l + 1];
    cpx<R> *t = prec<R>(L).t;
    level_apply<R>(c, L, z, t);
    tdz::launch_axpby_t<R>(c->st, one, r, mone, t, L.N);                   // t = r - A z
    transfer<R>(c, L.R, L.n, Cl.n, t, prec<R>(Cl).r, s1, s2, false);       // r_c = P^T t
    vcycle<R>(c, l + 1, prec<R>(Cl).r, prec<R>(Cl).z);
    transfer<R>(c, L.P, Cl.n, L.n, prec<R>(Cl).z, z, s1, s2, true);        // z += P z_c
    for (int s = 0; s < nu; ++s) smooth<R>(c, L, r, z, s1, s2, S.tmp[3], cheb);
}

// ---- exact Galerkin potential of the coarse levels
// Host pair table of level axis A against the finest level's Gauss points (axis F): for
// every (function i, offset d) the products N_i N_{i+d} at the run of fine points where
// both are nonzero (wall functions are not in the space: no entries).
struct PairTableHost {
    std::vector<int> q0, len, off;
    std::vector<double> val;
};

PairTableHost pair_table(const tdz::Axis1D &F, const tdz::Axis1D &A)
{
    PairTableHost t;
    if (!A.active) {   // unused axis: one function, one point of weight 1
        t.q0 = {0}; t.len = {1}; t.off = {0}; t.val = {1.0};
        return t;
    }
    const int p = A.p, n = A.n, D = 2 * p + 1, npts = F.E * F.Q;
    std::vector<std::vector<double>> runs((size_t)D * n);
    std::vector<int> start((size_t)D * n, -1);
    std::vector<double> v(p + 1);
    for (int m = 0; m < npts; ++m) {
        int f0 = 0;
        tdz::basis_values(A.t, p, F.X[m], f0, v.data());
        for (int a = 0; a <= p; ++a) {
            if (A.isWall[f0 + a]) continue;
            for (int b = 0; b <= p; ++b) {
                if (A.isWall[f0 + b]) continue;
                const size_t k = (size_t)(b - a + p) * n + (f0 + a);
                if (start[k] < 0) start[k] = m;
                std::vector<double> &r = runs[k];
                r.resize((size_t)(m - start[k]), 0.0);   // a gap (none for nested spaces) holds zeros
                r.push_back(v[a] * v[b]);
            }
        }
    }
    for (size_t k = 0; k < runs.size(); ++k) {
        t.q0.push_back(std::max(start[k], 0));
        t.len.push_back((int)runs[k].size());
        t.off.push_back((int)t.val.size());
        t.val.insert(t.val.end(), runs[k].begin(), runs[k].end());
    }
    return t;
}

// complex64 copies of the coarse potential stencils (mixed-precision V-cycle)
void galerkin_float(tdz_ctx *c)
{
    for (size_t l = 1; l < c->lv.size(); ++l) {
        Level &L = *c->lv[l];
        if (!L.Sv || L.Svf) continue;
        L.Svf = track(c, dalloc<float>(L.Sv_len));
        tdz::launch_stencil_demote(c->st, L.Sv, L.Svf, L.Sv_len);
    }
    HIPCHK(hipStreamSynchronize(c->st));
}

// Assemble P^T (s V) P for every coarse level once (the potential does not change in
// time; the field and the other terms coarsen per axis).
void ensure_galerkin(tdz_ctx *c)
{
    if (!c->ph.pot_kind || c->lv.size() < 2) return;
    Level &F = *c->lv[0];
    std::vector<double *> tmp;
    try {
        double *X[3], *W[3];
        for (int d = 0; d < 3; ++d) {
            X[d] = dalloc<double>(F.ax[d].X.size());
            tmp.push_back(X[d]);
            W[d] = dalloc<double>(F.ax[d].W.size());
            tmp.push_back(W[d]);
            HIPCHK(hipMemcpy(X[d], F.ax[d].X.data(), F.ax[d].X.size() * 8, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(W[d], F.ax[d].W.data(), F.ax[d].W.size() * 8, hipMemcpyHostToDevice));
        }
        for (size_t l = 1; l < c->lv.size(); ++l) {
            Level &L = *c->lv[l];
            if (L.Sv) continue;
            tdz::GalerkinPlan g;
            long len = 1, per_z = 0;
            for (int d = 0; d < 3; ++d) {
                const PairTableHost h = pair_table(F.ax[d], L.ax[d]);
                int *q0 = dalloc<int>(h.q0.size()), *ln = dalloc<int>(h.len.size()), *of = dalloc<int>(h.off.size());
                double *val = dalloc<double>(std::max<size_t>(h.val.size(), 1));
                tmp.insert(tmp.end(), {reinterpret_cast<double *>(q0), reinterpret_cast<double *>(ln),
                                       reinterpret_cast<double *>(of), val});
                HIPCHK(hipMemcpy(q0, h.q0.data(), h.q0.size() * 4, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(ln, h.len.data(), h.len.size() * 4, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(of, h.off.data(), h.off.size() * 4, hipMemcpyHostToDevice));
                if (!h.val.empty()) HIPCHK(hipMemcpy(val, h.val.data(), h.val.size() * 8, hipMemcpyHostToDevice));
                g.pairs[d] = {(int)h.q0.size(), q0, ln, of, val};
                g.X[d] = X[d];
                g.W[d] = W[d];
                g.nq[d] = F.ax[d].E * F.ax[d].Q;
                len *= (long)h.q0.size();
            }
            g.pot = {c->ph.pot_kind, c->ph.s, c->ph.pot[0], c->ph.pot[1]};
            per_z = (long)g.nq[1] * g.pairs[0].npairs + (long)g.pairs[1].npairs * g.pairs[0].npairs;
            const long work_len = per_z * std::min<long>(g.nq[2], std::max<long>(1, (128L << 20) / per_z));
            double *work = dalloc<double>(work_len);
            tmp.push_back(work);
            L.Sv = track(c, dalloc<double>(len));
            L.Sv_len = len;
            tdz::galerkin_potential(c->st, g, L.Sv, work, work_len);
            HIPCHK(hipStreamSynchronize(c->st));
        }
    } catch (...) {
        for (double *q : tmp) hipFree(q);
        throw;
    }
    for (double *q : tmp) HIPCHK(hipFree(q));
    if (c->mixed_ready) galerkin_float(c);
}

// Single-precision data for the mixed-precision preconditioners.
void setup_mixed(tdz_ctx *c)
{
    if (c->mixed_ready) return;
    const long N = c->lv[0]->N;
    for (int i = 0; i < 4; ++i) c->sf.tmp[i] = track(c, dalloc<fcomplex>(N));
    c->sf.w1 = track(c, dalloc<fcomplex>(N));
    c->sf.w2 = track(c, dalloc<fcomplex>(N));
    for (auto &L : c->lv) upload_level_tables_f(c, *L);
    c->mixed_ready = true;
    galerkin_float(c);
    set_shift_bands(c);
}

// Lumped non-separable potential at the Greville points (s V minus its separable part)
// and the correction bands, for the first-order corrected preconditioner.
void setup_corrected(tdz_ctx *c)
{
    if (c->dVg) return;
    Level &L = *c->lv[0];
    std::vector<double> gr[3];
    for (int d = 0; d < 3; ++d)
        gr[d] = d < c->prob.dim ? tdz::greville(L.ax[d].t, L.ax[d].p) : std::vector<double>{0.0};
    std::vector<float> vg(L.N, 0.f);
    if (c->ph.pot_kind == 2)
        for (int k = 0; k < L.n[2]; ++k)
            for (int j = 0; j < L.n[1]; ++j)
                for (int i = 0; i < L.n[0]; ++i) {
                    const double r2 = gr[0][i] * gr[0][i] + gr[1][j] * gr[1][j] + gr[2][k] * gr[2][k];
                    vg[i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k)] =
                        (float)(-c->ph.s * c->ph.pot[0] / std::sqrt(r2 + c->ph.pot[1]));
                }
    c->dVg = upload(c, vg);
    for (int d = 0; d < 3; ++d) c->dBcorr[d] = track(c, dalloc<fcomplex>(L.M[d].size()));
    const long N = L.N;
    c->sf.u1 = track(c, dalloc<fcomplex>(N));
    c->sf.u2 = track(c, dalloc<fcomplex>(N));
    c->sf.y0 = track(c, dalloc<fcomplex>(N));
    set_shift_bands(c);
}

// z ~ A^-1 r by the first-order corrected separable solve, complex64:
//   y = S^-1 r,  z = S^-1 (r - g Vg .* M y - g (CAP + sE.D) y - n(y))
// S = alpha M + g (K + V_sep) (fast diagonalisation); n(y) is the learned remainder
// of the potential product (null: lumped product only).
// Size of the terms the corrected solve treats as a perturbation of S, relative to
// S >= M: |g| (max W + sum_a |sE_a| max|x_a| + max|s V - V_sep|). The first-order
// correction only helps while this is well below 1.
double correction_bound(tdz_ctx *c)
{
    Level &L = *c->lv[0];
    double w = 0, f = 0;
    for (int d = 0; d < c->prob.dim; ++d) {
        double wm = 0;
        for (double v : L.ax[d].cap) wm = std::max(wm, std::fabs(v));
        if (c->cap_on) w += wm;
        f += std::fabs(c->sE[d]) * std::max(std::fabs(L.ax[d].lo), std::fabs(L.ax[d].hi));
    }
    double v = 0;
    if (c->ph.pot_kind == 2) v = std::fabs(c->ph.pot[0]) / std::sqrt(std::max(c->ph.pot[1], 1e-300));
    return std::abs(c->g) * (w + f + v);
}

// Bands of the fused products on the fine level (complex64 path).
tdz::FusedBands fused_bands(tdz_ctx *c)
{
    Level &L = *c->lv[0];
    LearnState &ls = c->learn;
    tdz::FusedBands fb;
    for (int d = 0; d < 3; ++d) {
        fb.n[d] = L.n[d];
        fb.p[d] = L.ax[d].p;
        fb.M[d] = L.f.M[d];
        fb.X[d] = ls.Xr[d];
        fb.C[d] = c->dBcorr[d];
        fb.g[d] = ls.gr[d];
    }
    return fb;
}

// One training sample from the complex64 vector y: the network input (y and its first
// moments, the fused products also leave M y in the level's t) and the target
// t = V y - Vg .* M y by the matrix-free complex64 potential kernel.
void harvest_sample(tdz_ctx *c, const fcomplex *y)
{
    Level &L = *c->lv[0];
    const long N = L.N;
    LearnState &ls = c->learn;
    constexpr int CI = tdz::kCnnInputs;
    fcomplex *tmp[7] = {c->sf.tmp[0], c->sf.tmp[1], c->sf.tmp[2], c->sf.tmp[3], c->sf.w1, c->sf.w2, c->sf.u2};
    tdz::fused_products(c->st, fused_bands(c), ls.sin, y, tmp, L.f.t, c->sf.u1, ls.feat);
    fcomplex *ys = track(c, dalloc<fcomplex>((size_t)CI * N)), *ts = track(c, dalloc<fcomplex>(N));
    HIPCHK(hipMemcpyAsync(ys, ls.feat, CI * N * sizeof(fcomplex), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemsetAsync(c->sf.u2, 0, N * sizeof(fcomplex), c->st));
    pot_apply<float>(c, L, 1.0, y, c->sf.u2);
    tdz::lumped_residual(c->st, N, c->sf.u2, L.f.t, c->dVg, ts);
    std::vector<fcomplex> h(N);
    HIPCHK(hipMemcpyAsync(h.data(), ts, N * sizeof(fcomplex), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double t2 = 0;
    for (const auto &v : h) t2 += (double)v.re * v.re + (double)v.im * v.im;
    ls.ys.push_back(ys);
    ls.ts.push_back(ts);
    ls.tnorm2.push_back(t2);
}

// Synthetic training samples that cover the whole grid and every scale: y = S^-1 w for a
// random field w (even samples: the separable solve's own spectrum, rough), or
// y = S^-1 M S^-1 w (odd samples: smoother). The network's target is a linear map of y,
// so inputs spread over the domain fit it where the harvested ones (near the initial
// state) do not reach.
void random_samples(tdz_ctx *c, int count, unsigned long long seed)
{
    Level &L = *c->lv[0];
    const long N = L.N;
    for (int i = 0; i < count; ++i) {
        tdz::random_field(c->st, N, seed + 7919ull * (unsigned long long)i, L.f.r);
        fd_apply<float>(c, L, L.f.r, L.f.z, c->sf.w1, c->sf.w2);
        if (i % 2) {
            tdz::SepBandsT<float> mb = sep_bands<float>(L);
            for (int d = 0; d < 3; ++d) mb.B[d] = L.f.Bmass[d];
            tdz::launch_sep_apply<float>(c->st, mb, L.f.z, L.f.r, c->sf.tmp);
            fd_apply<float>(c, L, L.f.r, L.f.z, c->sf.w1, c->sf.w2);
        }
        harvest_sample(c, L.f.z);
    }
}

// Sweeps (and, for the learned preconditioner, whether the network runs in the last one)
// for the next solve of the corrected preconditioners, by expected cost: k sweeps contract
// the residual by rho per Krylov iteration, so the solve needs about
// its = ceil(log(tol / r0) / log(rho)) iterations (tol the relative tolerance, r0 the
// solve's initial relative residual), each costing an operator apply and the Krylov
// work (about 2.1 units of 1 ms at 128^3) plus the preconditioner (0.13 units per separable
// solve, 0.062 per correction, 0.24 for the network). rho is the mean contraction per
// iteration, (final / initial residual)^(1/its), of the last solve that made that choice;
// an unobserved sweep count is estimated from
// an observed one of the same choice as rho_j = rho_k^((j+1)/(k+1)) (the error after k
// sweeps is (S^-1 dA)^(k+1)), and an unobserved network choice from the other one (the
// network assumed to gain half an order per iteration where only the network-free choice
// has been observed; the reverse is taken as equal). The first solve after pc_setup takes
// two sweeps (with the network); from then on the other network choice is tried at solve 4
// and then every 32 solves, and a solve that picks k > 1 tries k - 1 at solve 12 and then
// every 32 solves; either interval is doubled (up to 512) after each try that gained no
// iterations, so that estimates follow the field at little cost. Where a choice has been observed at
// a similar target (within a decade of tol / r0), the moving average of its iteration counts
// (weight 0.05 per solve) replaces the prediction from rho.
void choose_sweeps(tdz_ctx *c)
{
    constexpr int kMaxSweeps = 3;
    const bool can_net = c->pc_kind == TDZ_PC_LEARNED && c->learn.trained;
    c->pc_sweeps = 1;
    c->pc_net = can_net;
    c->pc_cform = false;
    if (c->pc_r0rel <= 0) return;
    const double target = c->pc_rtol / c->pc_r0rel;
    // the correction form's single-FP16 corrections leave about 1e-4 of the residual in a
    // strong field: right for solves of two or more iterations, not for one
    c->pc_cform = target < 1e-6;
    auto observed_ref = [&](int net) {
        int ref = -1;
        for (int k = 1; k <= kMaxSweeps; ++k)
            if (c->rho_hat[net][k] > 0 && (ref < 0 || k < ref)) ref = k;
        return ref;
    };
    auto estimate = [&](int net, int k) {
        if (c->rho_hat[net][k] > 0) return c->rho_hat[net][k];
        const int ref = observed_ref(net);
        if (ref > 0) return std::pow(c->rho_hat[net][ref], (k + 1.0) / (ref + 1.0));
        const int other = 1 - net, oref = observed_ref(other);
        const double po = net ? 1.5 : 1.0;   // the unobserved network gains half an order
        if (oref > 0) {
            const double ro = c->rho_hat[other][k] > 0 ? c->rho_hat[other][k]
                                                       : std::pow(c->rho_hat[other][oref], (k + 1.0) / (oref + 1.0));
            return std::pow(ro, po);
        }
        return std::pow(1e-2, (k + 1.0) * (net ? 1.5 : 1.0));
    };
    bool seen = false;
    for (const auto &row : c->rho_hat)
        for (double v : row) seen = seen || v > 0;
    if (!seen) {   // the first solve after pc_setup: two sweeps, with the network if there is one
        ++c->pc_solves;
        c->pc_exploring = c->pc_flipping = false;
        c->pc_sweeps = 2;
        ++c->sweep_count[can_net ? 1 : 0][2];
        return;
    }
    int best = 1, best_net = can_net ? 1 : 0;
    double best_cost = 1e300;
    for (int net = can_net ? 1 : 0; net >= 0; --net)
        for (int k = 1; k <= kMaxSweeps; ++k) {
            const double rho = std::min(estimate(net, k), 0.5);
            // a solve that met the tolerance at its contraction is counted as such (-0.05: the
            // observed mean contraction of an its-iteration solve reproduces its iterations)
            double its =
                target >= 1.0 ? 1.0 : std::max(1.0, std::ceil(std::log(target) / std::log(rho) - 0.05));
            // near the targets it has seen, a choice's own iteration average decides: one
            // harder solve then moves the estimate by a twentieth of an iteration, not a whole one
            if (c->its_hat[net][k] > 0 && std::fabs(std::log10(target) - c->its_lt[net][k]) < 1.0)
                its = c->its_hat[net][k];
            // ms at 128^3: operator apply and Krylov work per iteration, a complex64 separable
            // solve, a fused correction pass, the network (its feature pass, two layers, combine)
            const double cost = its * (2.1 + 0.13 * (k + 1) + 0.062 * k + (net ? 0.24 : 0.0));
            if (cost < best_cost * (1 - 1e-9)) {
                best_cost = cost;
                best = k;
                best_net = net;
            }
        }
    ++c->pc_solves;
    c->pc_exploring = false;
    if (best > 1 && c->pc_solves >= c->pc_next_drop) {
        --best;
        c->pc_exploring = true;
        c->pc_next_drop = c->pc_solves + c->pc_drop_iv;
    }
    c->pc_flipping = false;
    if (can_net && c->pc_solves >= c->pc_next_flip) {
        c->pc_flipping = true;
        c->pc_flip_from_k = best;
        c->pc_flip_from_net = best_net == 1;
        best_net = 1 - best_net;
        c->pc_next_flip = c->pc_solves + c->pc_flip_iv;
    }
    c->pc_sweeps = best;
    c->pc_net = best_net == 1;
    ++c->sweep_count[c->pc_net ? 1 : 0][best];
}

// the current preconditioner is a corrected solve with adaptive sweeps
bool sweeps_apply(const tdz_ctx *c)
{
    return c->pc_kind == TDZ_PC_CORRECTED || c->pc_kind == TDZ_PC_CORRECTED_EXACT ||
           (c->pc_kind == TDZ_PC_LEARNED && !c->learn.fallen);
}

// Record the mean contraction per iteration of a solve that used k sweeps (with or without
// the network); counts below k are at least as bad as the model says.
void observe_sweeps(tdz_ctx *c, int k, bool net, double rho)
{
    if (k < 1 || k > 3 || !(rho > 0)) return;
    double *r = c->rho_hat[net ? 1 : 0];
    r[k] = rho;
    for (int j = 1; j < k; ++j) r[j] = std::max(r[j], std::pow(rho, (j + 1.0) / (k + 1.0)));
}

void corrected_apply(tdz_ctx *c, const dcomplex *r, dcomplex *z, bool exact_v = false, bool use_net = false)
{
    Level &L = *c->lv[0];
    const long N = L.N;
    if (c->pc_choose) {   // sweeps, network and form of this solve
        choose_sweeps(c);
        c->pc_choose = false;
    }
    tdz::launch_demote(c->st, r, L.f.r, N);
    // Correction form (corr_form 1, with the split-FP16 transforms): y = S^-1 r is kept, and
    // each sweep solves for the correction alone, z = y + S^-1 (-dA~ z), whose right-hand side
    // is about |S^-1 dA| ~ 1e-2 of r, so its transforms run on single FP16 products (about
    // 1e-3 relative, 1e-5 of z) at 80% of the split form's cost. Adaptive solves take it
    // when they need two or more Krylov iterations (tol / r0 < 1e-6, choose_sweeps).
    const bool cform = c->corr_form == 1 && c->fd_transform > 0 && (c->corrections > 0 || c->pc_cform);
    fcomplex *y0 = cform ? c->sf.y0 : L.f.z;
    const int ckind = cform ? 3 : -1;   // single FP16 products
    fd_apply<float>(c, L, L.f.r, y0, c->sf.w1, c->sf.w2);                // y = S^-1 r
    if (correction_bound(c) > 0.25) {   // perturbation too large: the separable solve alone
        tdz::launch_promote(c->st, y0, z, N);
        return;
    }
    LearnState &ls = c->learn;
    const int sweeps = c->corrections > 0 ? c->corrections : c->pc_sweeps;
    // z <- S^-1 (r - dA~ z), repeated: the error after k corrections is (S^-1 dA)^(k+1)
    // plus, at first order, what the last sweep's dA~ misses of dA (the lumped potential:
    // ~1e-4 of r at dt = 0.02; with the network's remainder far less); an earlier sweep's
    // miss enters multiplied by S^-1 dA, so the network runs in the last sweep only
    const fcomplex *zc = y0;                        // the current iterate
    const fcomplex *rr = cform ? nullptr : L.f.r;   // the sweeps' right-hand side: r, or zero
    bool promoted = false;
    // the sweep's solve into L.f.z; in the correction form y is added (into z on the last sweep)
    auto solve = [&](bool last) {
        fd_apply<float>(c, L, c->sf.u1, L.f.z, c->sf.w1, c->sf.w2, ckind);
        if (cform) {
            if (last) {
                tdz::launch_promote_add(c->st, y0, L.f.z, z, N);
                promoted = true;
            } else {
                tdz::launch_axpby_t<float>(c->st, {1.f, 0.f}, y0, {1.f, 0.f}, L.f.z, N);
            }
        }
        zc = L.f.z;
    };
    for (int k = 0; k < sweeps; ++k) {
        const bool net = use_net && ls.trained && k == sweeps - 1 && (c->corrections > 0 || c->pc_net);
        const bool harvest = k == 0 && ls.harvesting && (int)ls.ys.size() < ls.max_samples && c->ph.pot_kind;
        // a sweep without the network: its right-hand side r - dA~ z (or - dA~ z) in one
        // z-marching pass
        if (!harvest && !net && !exact_v &&
            tdz::fused_correction(c->st, fused_bands(c), zc, rr, c->dVg, cx<float>(c->g), c->sf.u1, c->sf.tmp[0],
                                  c->sf.tmp[1])) {
            solve(k == sweeps - 1);
            continue;
        }
        if (harvest || net || ls.ready) {
            // M z, g (CAP + sE.D) z and, for the network, its input, in one z-marching pass
            // (equal orders in 3D; three passes otherwise)
            fcomplex *tmp[7] = {c->sf.tmp[0], c->sf.tmp[1], c->sf.tmp[2], c->sf.tmp[3], c->sf.w1, c->sf.w2, c->sf.u2};
            tdz::fused_products(c->st, fused_bands(c), ls.sin, zc, tmp, L.f.t, c->sf.u1,
                                harvest || net ? ls.feat : nullptr);
        } else {
            tdz::SepBandsT<float> mb = sep_bands<float>(L);
            for (int d = 0; d < 3; ++d) mb.B[d] = L.f.Bmass[d];
            tdz::launch_sep_apply<float>(c->st, mb, zc, L.f.t, c->sf.tmp);       // M z
            tdz::SepBandsT<float> cb = sep_bands<float>(L);
            for (int d = 0; d < 3; ++d) cb.B[d] = c->dBcorr[d];
            tdz::launch_sep_apply<float>(c->st, cb, zc, c->sf.u1, c->sf.tmp);    // g (CAP + sE.D) z
        }
        const fcomplex *n = nullptr;
        if (exact_v && c->ph.pot_kind) {   // potential product by the matrix-free kernel (reference for the learned one)
            HIPCHK(hipMemsetAsync(c->sf.u2, 0, N * sizeof(fcomplex), c->st));
            pot_apply<float>(c, L, c->g, zc, c->sf.u2);
            HIPCHK(hipMemsetAsync(L.f.t, 0, N * sizeof(fcomplex), c->st));
            n = c->sf.u2;
        }
        if (harvest) harvest_sample(c, zc);
        const float *no = nullptr;   // the network's output, scaled by g / sout in the combine
        if (net) {
            const int n3[3] = {L.n[0], L.n[1], L.n[2]};
            // the weights' BF16 operand packing persists in the work buffer until they change
            tdz::cnn_infer_mfma(c->st, n3, ls.feat, ls.W1, ls.W2, ls.fgen, ls.mb, ls.out, fcomplex{0.f, 0.f}, nullptr,
                                ls.packed_ver != ls.wver);
            ls.packed_ver = ls.wver;
            no = ls.out;
        }
        tdz::launch_corr_combine(c->st, rr, L.f.t, c->sf.u1, n, c->dVg, cx<float>(c->g), c->sf.u1, N, no,
                                 cx<float>(c->g / ls.sout));
        solve(k == sweeps - 1);
    }
    if (!promoted) tdz::launch_promote(c->st, zc, z, N);
}

void pc_apply_dev(tdz_ctx *c, const dcomplex *r, dcomplex *z)
{
    Level &L = *c->lv[0];
    switch (c->pc_kind) {
    case TDZ_PC_SEPARABLE: fd_apply<double>(c, L, r, z, c->sd.w1, c->sd.w2); break;
    case TDZ_PC_MG:
        refresh_coarse(c);
        vcycle<double>(c, 0, r, z);
        break;
    case TDZ_PC_SEPARABLE_MIXED:
        tdz::launch_demote(c->st, r, L.f.r, L.N);
        fd_apply<float>(c, L, L.f.r, L.f.z, c->sf.w1, c->sf.w2);
        tdz::launch_promote(c->st, L.f.z, z, L.N);
        break;
    case TDZ_PC_CORRECTED: corrected_apply(c, r, z); break;
    case TDZ_PC_CORRECTED_EXACT: corrected_apply(c, r, z, true); break;
    case TDZ_PC_LEARNED:
        if (c->learn.fallen) {
            if (c->learn.fallback_kind == TDZ_PC_MG) {
                refresh_coarse(c);
                vcycle<double>(c, 0, r, z);
            }
            else fd_apply<double>(c, L, r, z, c->sd.w1, c->sd.w2);
        } else {
            corrected_apply(c, r, z, false, true);
        }
        break;
    case TDZ_PC_MG_MIXED:
        tdz::launch_demote(c->st, r, L.f.r, L.N);
        refresh_coarse(c);
        vcycle<float>(c, 0, L.f.r, L.f.z);
        tdz::launch_promote(c->st, L.f.z, z, L.N);
        break;
    default: HIPCHK(hipMemcpyAsync(z, r, L.N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st)); break;
    }
}

// <x, y> on device (deterministic), result on host
cplx dev_dot(tdz_ctx *c, const dcomplex *x, const dcomplex *y, long N)
{
    tdz::launch_multi_dot(c->st, x, N, 1, y, N, c->dpart, c->dh);
    dcomplex h{0, 0};
    HIPCHK(hipMemcpyAsync(&h, c->dh, sizeof(dcomplex), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return {h.re, h.im};
}

// y = M x on the fine level (walls: identity)
void mass_apply(tdz_ctx *c, const dcomplex *x, dcomplex *y)
{
    Level &L = *c->lv[0];
    tdz::SepBandsT<double> mb = sep_bands<double>(L);
    for (int d = 0; d < 3; ++d) mb.B[d] = L.d.Bmass[d];
    tdz::launch_sep_apply<double>(c->st, mb, x, y, c->sd.tmp);
    if (c->prob.dirichlet) {
        const unsigned char *w[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
        tdz::launch_wall_fix<double>(c->st, L.n, w, x, y);
    }
}

double dev_norm2(tdz_ctx *c, const dcomplex *x, long N)
{
    tdz::launch_norm2(c->st, x, N, c->dpartr, c->dnorm);
    double h = 0;
    HIPCHK(hipMemcpyAsync(&h, c->dnorm, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return h;
}

// FGMRES(m), right preconditioning, classical Gram-Schmidt with one
// re-orthogonalisation (two fused passes), complex Givens rotations. Stops when
// the residual estimate |g_{j+1}| <= rtol ||b|| (true residual for right PC).
// With r0_in_v0 the caller has already stored the initial residual b - A x in V_0.
// Returns the iteration count; without convergence in maxit iterations it throws, or
// with allow_fail returns -1 (x then holds the last iterate).
// With ax_out, a converged solve also returns A x = b - r there, r the final residual
// taken from the Arnoldi relation of the last cycle: r = V_{j+1} (beta e_1 - Hbar y)
// = gamma_{j+1} V_{j+1} Q^H e_{j+1} (Q the Givens rotations), at the cost of one
// (j+1)-term vector update instead of an operator apply.
// The host waits for the device once per Arnoldi step (the projections) and once at the
// start (||b|| and, with r0_in_v0, ||r0|| together); the solution update at the end is
// queued without a wait (its coefficients travel in the kernel arguments; above
// kCoefArgMax they are uploaded from pinned staging, which the next solve rewrites only
// after its own first wait).
int fgmres(tdz_ctx *c, const dcomplex *b, dcomplex *x, double rtol, int maxit, tdz_ksp_stats *st,
           bool r0_in_v0 = false, bool allow_fail = false, dcomplex *ax_out = nullptr)
{
    Level &L = *c->lv[0];
    const long N = L.N;
    const int m = c->restart;
    auto Vj = [&](int j) { return c->Vb + (long)j * N; };
    auto Zj = [&](int j) { return c->pc_kind == TDZ_PC_NONE ? Vj(j) : c->Zb + (long)j * N; };
    std::vector<cplx> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), hv(m + 1), hv2(m + 1);
    double bnorm, beta0 = -1;   // beta0 >= 0: ||V_0|| of the first cycle, fetched with ||b||
    // pinned staging: [0, 2(m+1)] the projections and the norm read back (the device layout of
    // dh, dh2, dnorm); then the solution and residual coefficients of the update (passed in
    // the kernel arguments, or uploaded to dh, dh2 from here for restarts above kCoefArgMax)
    const int slen = 4 * (m + 1) + 1;
    if (c->hstage_len < slen) {
        HIPCHK(hipStreamSynchronize(c->st));   // no queued upload still reads the old staging
        if (c->hstage) HIPCHK(hipHostFree(c->hstage));
        c->hstage = nullptr;
        HIPCHK(hipHostMalloc((void **)&c->hstage, slen * sizeof(dcomplex), hipHostMallocDefault));
        c->hstage_len = slen;
    }
    const bool queued = c->norms_queued;
    c->norms_queued = false;
    if (r0_in_v0) {
        if (!queued) {
            tdz::launch_norm2(c->st, b, N, c->dpartr, c->dnorm + 1);
            tdz::launch_norm2(c->st, Vj(0), N, c->dpartr, c->dnorm);
        }
        // into the readback part of the staging (an earlier solve's queued upload, if any,
        // reads the coefficient part, and precedes this copy on the stream)
        double *h2 = reinterpret_cast<double *>(c->hstage);
        HIPCHK(hipMemcpyAsync(h2, c->dnorm, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        bnorm = std::sqrt(h2[1]);
        beta0 = std::sqrt(h2[0]);
    } else {
        bnorm = std::sqrt(dev_norm2(c, b, N));
    }
    const dcomplex *hback = c->hstage;
    dcomplex *ystage = c->hstage + 2 * (m + 1) + 1, *rstage = ystage + (m + 1);
    double last_hn = 0;        // ||w|| of the last Arnoldi step (after both Gram-Schmidt passes)
    bool last_fused = false;   // its second pass deferred: w as stored still holds V h2
    std::vector<dcomplex> last_h2(tdz::kCgsFusedMax);
    int its = 0, restarts = 0;
    double res = 0;
    bool chose = false;   // this solve's sweeps were chosen by the cost model
    if (bnorm == 0.0) {
        HIPCHK(hipMemsetAsync(x, 0, N * sizeof(dcomplex), c->st));
        if (ax_out) HIPCHK(hipMemsetAsync(ax_out, 0, N * sizeof(dcomplex), c->st));
        if (st) *st = {0, 0, 0.0, 0.0, c->rebuilds};
        return 0;
    }
    bool converged = false;
    for (;;) {
        if (!r0_in_v0) {
            level_apply(c, L, x, Vj(0));
            tdz::launch_axpby(c->st, {1, 0}, b, {-1, 0}, Vj(0), N);   // V0 = b - A x
        }
        r0_in_v0 = false;
        const double beta = beta0 >= 0 ? beta0 : std::sqrt(dev_norm2(c, Vj(0), N));
        beta0 = -1;
        res = beta / bnorm;
        if (res <= rtol) {
            converged = true;
            if (ax_out) {   // V_0 holds the residual b - A x itself
                tdz::LinComb ax;
                ax.add({1, 0}, b);
                ax.add({-1, 0}, Vj(0));
                tdz::launch_lincomb(c->st, ax, ax_out, N);
            }
            break;
        }
        if (its >= maxit) break;
        const bool adapt = restarts == 0 && c->corrections == 0 && sweeps_apply(c);
        if (adapt) {   // the corrected preconditioners pick their sweeps from r0 and rtol
            c->pc_r0rel = res;
            c->pc_rtol = rtol;
            c->pc_choose = true;
            chose = true;
        }
        ++restarts;
        tdz::launch_axpby(c->st, {1.0 / beta, 0}, Vj(0), {0, 0}, Vj(0), N);
        std::fill(gv.begin(), gv.end(), cplx(0));
        gv[0] = beta;
        int j = 0;
        for (; j < m && its < maxit; ++j) {
            if (c->pc_kind != TDZ_PC_NONE) pc_apply_dev(c, Vj(j), Zj(j));
            dcomplex *w = Vj(j + 1);
            level_apply(c, L, Zj(j), w);
            // CGS2: two passes of h = V^H w, w -= V h. Up to kCgsFusedMax vectors the first
            // pass also takes the second pass's projections h2 = V^H w1 and ||w1||^2; the
            // second subtraction is deferred (||w1 - V h2||^2 = ||w1||^2 - ||h2||^2 for
            // orthonormal V) into the normalisation, or into the residual coefficients of a
            // converged solve
            const int nv = j + 1;
            const bool fused_cgs = nv <= tdz::kCgsFusedMax;
            tdz::launch_multi_dot(c->st, c->Vb, N, nv, w, N, c->dpart, c->dh);
            if (fused_cgs) {
                tdz::launch_multi_axpy_norm_dot(c->st, c->Vb, N, nv, c->dh, w, N, c->dpartr, c->dpart, c->dnorm,
                                                c->dh2);
            } else {
                tdz::launch_multi_axpy_norm(c->st, c->Vb, N, nv, c->dh, w, N, c->dpartr, c->dnorm);
                tdz::launch_multi_dot(c->st, c->Vb, N, nv, w, N, c->dpart, c->dh2);
                tdz::launch_multi_axpy_norm(c->st, c->Vb, N, nv, c->dh2, w, N, c->dpartr, c->dnorm);
            }
            HIPCHK(hipMemcpyAsync(c->hstage, c->dh, (2 * (m + 1) + 1) * sizeof(dcomplex), hipMemcpyDeviceToHost,
                                  c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            const dcomplex *h1 = hback, *h2 = hback + (m + 1);
            const double wn2 = reinterpret_cast<const double *>(hback + 2 * (m + 1))[0];
            for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = cplx(h1[i].re + h2[i].re, h1[i].im + h2[i].im);
            double hn2 = wn2;
            bool deferred = fused_cgs;
            if (fused_cgs) {
                double h2sq = 0;
                for (int i = 0; i <= j; ++i) h2sq += h2[i].re * h2[i].re + h2[i].im * h2[i].im;
                if (h2sq > 0.5 * wn2) {
                    // w1 lies largely in span(V): ||w1||^2 - ||h2||^2 would cancel, so the
                    // second pass is applied now and the norm taken of its result
                    tdz::launch_multi_axpy_norm(c->st, c->Vb, N, nv, c->dh2, w, N, c->dpartr, c->dnorm);
                    HIPCHK(hipMemcpyAsync(&hn2, c->dnorm, sizeof(double), hipMemcpyDeviceToHost, c->st));
                    HIPCHK(hipStreamSynchronize(c->st));
                    deferred = false;
                } else {
                    hn2 = wn2 - h2sq;
                }
            }
            const double hn = hn2 > 0 ? std::sqrt(hn2) : 0.0;
            H[(size_t)(j + 1) * m + j] = hn;
            last_hn = hn;
            last_fused = deferred;
            if (deferred) std::copy(h2, h2 + nv, last_h2.begin());
            for (int i = 0; i < j; ++i) {
                const cplx a = H[(size_t)i * m + j], bb = H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j] = cs[i] * a + sn[i] * bb;
                H[(size_t)(i + 1) * m + j] = -std::conj(sn[i]) * a + cs[i] * bb;
            }
            const cplx h1j = H[(size_t)j * m + j], h2j = H[(size_t)(j + 1) * m + j];
            const double rho = std::sqrt(std::norm(h1j) + std::norm(h2j));
            if (std::abs(h1j) == 0.0) { cs[j] = 0.0; sn[j] = 1.0; }
            else { cs[j] = std::abs(h1j) / rho; sn[j] = (h1j / std::abs(h1j)) * std::conj(h2j) / rho; }
            H[(size_t)j * m + j] = cs[j] * h1j + sn[j] * h2j;
            H[(size_t)(j + 1) * m + j] = 0.0;
            gv[j + 1] = -std::conj(sn[j]) * gv[j];
            gv[j] = cs[j] * gv[j];
            ++its;
            res = std::abs(gv[j + 1]) / bnorm;
            if (res <= rtol || hn == 0.0) { ++j; converged = true; break; }
            // V_{j+1} = (w - V h2) / ||w - V h2|| for the next step (a converged solve leaves w
            // as it is: only the residual update below reads it, with its coefficients adjusted)
            if (deferred) tdz::launch_correct_scale(c->st, c->Vb, N, nv, c->dh2, w, N, 1.0 / hn);
            else tdz::launch_scale_invsqrt(c->st, w, c->dnorm, w, N);
        }
        std::vector<cplx> yv(j);
        for (int i = j - 1; i >= 0; --i) {
            cplx s = gv[i];
            for (int k = i + 1; k < j; ++k) s -= H[(size_t)i * m + k] * yv[k];
            yv[i] = s / H[(size_t)i * m + i];
        }
        for (int i = 0; i < j; ++i) ystage[i] = dc(yv[i]);
        const bool with_ax = converged && ax_out;
        if (with_ax) {
            // q = Q^H e_j: the inverse rotations, newest first
            std::vector<cplx> q(j + 1, cplx(0));
            q[j] = 1.0;
            for (int i = j - 1; i >= 0; --i) {
                const cplx a = q[i], bb = q[i + 1];
                q[i] = cs[i] * a - sn[i] * bb;
                q[i + 1] = std::conj(sn[i]) * a + cs[i] * bb;
            }
            for (int i = 0; i <= j; ++i) rstage[i] = dc(-gv[j] * q[i]);
            if (last_hn > 0) {   // V_j = (w - [V h2]) / ||.|| with w as stored
                const cplx cw = cplx(rstage[j].re, rstage[j].im) / last_hn;
                if (last_fused)
                    for (int i = 0; i < j; ++i)
                        rstage[i] = dc(cplx(rstage[i].re, rstage[i].im) - cw * cplx(last_h2[i].re, last_h2[i].im));
                rstage[j] = dc(cw);
            }
        }
        const dcomplex *Zb = c->pc_kind == TDZ_PC_NONE ? c->Vb : c->Zb;
        if (j + 1 <= tdz::kCoefArgMax) {   // the coefficients in the kernel arguments: no upload
            tdz::launch_multi_axpy_args(c->st, Zb, N, j, ystage, x, N);
            if (with_ax) tdz::launch_multi_axpy_args(c->st, c->Vb, N, j + 1, rstage, ax_out, N, b);   // b + sum
        } else {   // y to dh and, with A x, the residual coefficients to dh2 = dh + m + 1: one copy
            HIPCHK(hipMemcpyAsync(c->dh, ystage, (with_ax ? (m + 1) + (j + 1) : j) * sizeof(dcomplex),
                                  hipMemcpyHostToDevice, c->st));
            tdz::launch_multi_axpy(c->st, Zb, N, j, c->dh, x, N);
            if (with_ax) tdz::launch_multi_axpy(c->st, c->Vb, N, j + 1, c->dh2, ax_out, N, b);   // ax_out = b + sum
        }
        if (converged || its >= maxit) break;
    }
    if (chose && converged && its > 0 && c->pc_r0rel > 0 && res > 0) {   // mean contraction per iteration
        observe_sweeps(c, c->pc_sweeps, c->pc_net, std::pow(std::min(res / c->pc_r0rel, 1.0), 1.0 / its));
        const int k = c->pc_sweeps, n = c->pc_net ? 1 : 0;
        if (c->pc_exploring && k + 1 <= 3) {   // one sweep fewer: try again soon only if it paid
            const double ih = c->its_hat[n][k + 1];
            c->pc_drop_iv = (ih > 0 && its > ih + 0.5) ? std::min(2 * c->pc_drop_iv, 512) : 32;
        }
        if (c->pc_flipping) {   // the other network choice: try again soon only if it saved iterations
            const double ih = c->its_hat[c->pc_flip_from_net ? 1 : 0][c->pc_flip_from_k];
            c->pc_flip_iv = (ih > 0 && its > ih - 0.5) ? std::min(2 * c->pc_flip_iv, 512) : 32;
        }
        if (k >= 1 && k <= 3) {
            const double lt = std::log10(c->pc_rtol / c->pc_r0rel);
            double &ih = c->its_hat[n][k], &il = c->its_lt[n][k];
            if (ih > 0 && std::fabs(lt - il) < 1.0) {
                ih = 0.95 * ih + 0.05 * its;
                il = 0.95 * il + 0.05 * lt;
            } else {
                ih = its;
                il = lt;
            }
        }
    }
    if (st) *st = {its, restarts, res, bnorm, c->rebuilds};
    if (!converged) {
        if (allow_fail) return -1;
        char msg[160];
        std::snprintf(msg, sizeof msg, "KSP did not converge: %d iterations, relative residual %.3e", its, res);
        throw Error(msg);
    }
    return its;
}

void set_shift(tdz_ctx *c, cplx alpha, cplx g, const double sE[3])
{
    if (c->layered && sE && (sE[0] != 0.0 || sE[1] != 0.0))
        throw Error("a layered structure takes fields along z only");
    c->alpha = alpha;
    c->g = g;
    for (int a = 0; a < 3; ++a) c->sE[a] = sE ? sE[a] : 0.0;
    ++c->shift_ver;
    set_shift_bands(c);
}

void cheb_setup(tdz_ctx *c)
{
    if (c->layered) throw Error("the Jacobi-Chebyshev smoother does not take layered structures");
    for (size_t l = 0; l + 1 < c->lv.size(); ++l)
        if (c->lv[l]->cheb_max <= 0) c->lv[l]->cheb_max = cheb_bound(c, *c->lv[l]);
}

void pc_setup(tdz_ctx *c, int kind)
{
    if (kind < TDZ_PC_NONE || kind > TDZ_PC_LEARNED) throw Error("unknown preconditioner kind");
    const bool mg = kind == TDZ_PC_MG || kind == TDZ_PC_MG_MIXED;
    if (mg && c->lv.size() < 2) throw Error("multigrid needs at least two levels");
    if (c->layered && (kind == TDZ_PC_CORRECTED || kind == TDZ_PC_CORRECTED_EXACT || kind == TDZ_PC_LEARNED))
        throw Error("a layered structure is separable: its separable solve is exact (no correction to learn)");
    c->pc_kind = kind;
    for (auto &row : c->rho_hat) std::fill(std::begin(row), std::end(row), -1.0);   // new contraction estimates
    for (auto &row : c->sweep_count) std::fill(std::begin(row), std::end(row), 0L);
    c->pc_solves = 0;
    c->pc_next_drop = 12;
    c->pc_drop_iv = 32;
    c->pc_exploring = false;
    c->pc_next_flip = 4;
    c->pc_flip_iv = 32;
    c->pc_flipping = false;
    for (auto &row : c->its_hat) std::fill(std::begin(row), std::end(row), 0.0);
    c->pc_sweeps = 1;
    c->pc_net = true;
    if (mg || (kind == TDZ_PC_LEARNED && c->lv.size() >= 2)) ensure_galerkin(c);
    if (mg)
        for (auto &L : c->lv) setup_fd(c, *L);
    if ((mg || kind == TDZ_PC_LEARNED) && c->smoother == 1) cheb_setup(c);
    else if (kind != TDZ_PC_NONE)
        setup_fd(c, *c->lv[0]);
    const bool corr = kind == TDZ_PC_CORRECTED || kind == TDZ_PC_CORRECTED_EXACT || kind == TDZ_PC_LEARNED;
    if (kind == TDZ_PC_LEARNED) {
        if (!c->learn.trained) throw Error("the learned preconditioner has not been trained (tdz_learn_train)");
        c->learn.fallback_kind = c->lv.size() >= 2 ? TDZ_PC_MG : TDZ_PC_SEPARABLE;
        for (auto &L : c->lv) setup_fd(c, *L);
    }
    if (kind == TDZ_PC_SEPARABLE_MIXED || kind == TDZ_PC_MG_MIXED || corr) setup_mixed(c);
    if (corr) setup_corrected(c);
    c->pc_ready = true;
}

double pulse_field(const tdz_pulse &pl, int a, double t)
{
    if (pl.omega <= 0 || pl.ncycles <= 0) return 0.0;
    const double T = pl.ncycles * 2.0 * M_PI / pl.omega;
    if (t < 0 || t > T) return 0.0;
    const double s = std::sin(M_PI * t / T);
    return pl.E0[a] * s * s * std::sin(pl.omega * t + pl.cep);
}

// Bands of the layered z axis from its tables: Mw = int (1/m) B_i B_j, Sw = int (1/m)
// B_i' B_j', Vl = int V_l B_i B_j, with the layer of each Gauss point (the BenDaniel-Duke
// kinetic term 1/2 grad.(1/m) grad and the band offsets; H semantics of
// src/assembly/sumfact.cpp:496-502 with cK = w/(2m)).
void layer_bands(const tdz_ctx *c, const tdz::Axis1D &A, std::vector<double> &Mw, std::vector<double> &Sw,
                 std::vector<double> &Vl)
{
    const int p = A.p, nen = p + 1, bw = 2 * p + 1, Q = A.Q;
    Mw.assign((size_t)A.n * bw, 0.0);
    Sw = Mw;
    Vl = Mw;
    for (int e = 0; e < A.E; ++e)
        for (int q = 0; q < Q; ++q) {
            const double x = A.X[(size_t)e * Q + q], w = A.W[(size_t)e * Q + q];
            int l = 0;
            while (l + 1 < (int)c->lay_m.size() && x >= c->lay_b[l + 1]) ++l;
            const double im = 1.0 / c->lay_m[l], v = c->lay_v[l];
            for (int a = 0; a < nen; ++a)
                for (int b = 0; b < nen; ++b) {
                    const size_t k = (size_t)(A.off[e] + a) * bw + (b - a + p);
                    const double Na = A.B0[((size_t)e * nen + a) * Q + q], Nb = A.B0[((size_t)e * nen + b) * Q + q];
                    const double Da = A.B1[((size_t)e * nen + a) * Q + q], Db = A.B1[((size_t)e * nen + b) * Q + q];
                    Mw[k] += w * im * Na * Nb;
                    Sw[k] += w * im * Da * Db;
                    Vl[k] += w * v * Na * Nb;
                }
        }
}

// Knot vectors of the multigrid levels, finest first (the hierarchy of
// src/multigrid.cpp:239-268). A level coarsens every axis that can still lose element
// boundaries and keep at least p+1 elements; the list ends when no axis coarsens, at
// mg_levels levels, or, for the automatic depth (mg_levels = 0), after the first coarse
// level with fewer than 2000 functions.
std::vector<std::array<std::vector<double>, 3>> level_knots(const std::array<std::vector<double>, 3> &finest, int dim,
                                                            int p, int mg_levels)
{
    std::vector<std::array<std::vector<double>, 3>> out{finest};
    const size_t depth = mg_levels > 0 ? (size_t)mg_levels : 20;
    auto functions = [&](const std::array<std::vector<double>, 3> &K) {
        long n = 1;
        for (int d = 0; d < dim; ++d) n *= (long)K[d].size() - p - 1;
        return n;
    };
    while (out.size() < depth) {
        if (mg_levels == 0 && out.size() > 1 && functions(out.back()) < 2000) break;
        std::array<std::vector<double>, 3> next = out.back();
        int coarsened = 0;
        for (int d = 0; d < dim; ++d) {
            std::vector<double> c;
            if (tdz::coarsen_knots(out.back()[d], c) && tdz::knot_spans(c) > p) {
                next[d].swap(c);
                ++coarsened;
            }
        }
        if (coarsened == 0) break;
        out.push_back(std::move(next));
    }
    return out;
}

void build(tdz_ctx *c, const tdz_problem &pr)
{
    if (pr.dim < 1 || pr.dim > 3) throw Error("dim must be 1, 2 or 3");
    if (pr.p < 1 || pr.p > 9) throw Error("spline degree must be in [1, 9]");
    if (pr.nquad < 1 || pr.nquad > 16) throw Error("nquad must be in [1, 16]");
    if (pr.mass <= 0) throw Error("mass must be positive");
    for (int d = 0; d < pr.dim; ++d) {
        if (pr.nelem[d] < 1) throw Error("nelem must be positive");
        if (!(pr.hi[d] > pr.lo[d])) throw Error("empty domain");
    }
    c->prob = pr;
    HIPCHK(hipGetDevice(&c->device));
    c->ph.mass = pr.mass;
    c->ph.s = -pr.charge;
    c->ph.pot_kind = pr.pot_kind;
    c->ph.cap_kind = pr.cap_kind;
    for (int i = 0; i < 4; ++i) { c->ph.pot[i] = pr.pot[i]; c->ph.cap[i] = pr.cap[i]; }
    HIPCHK(hipStreamCreate(&c->st));

    // ---- level hierarchy
    std::vector<std::array<std::vector<double>, 3>> knots;
    std::array<std::vector<double>, 3> k0;
    for (int d = 0; d < pr.dim; ++d) k0[d] = tdz::uniform_knots(pr.nelem[d], pr.lo[d], pr.hi[d], pr.p);
    if (pr.knots && pr.nknots > 0) {   // the z axis of a layered structure (last active axis)
        std::vector<double> t(pr.knots, pr.knots + pr.nknots);
        const int p = pr.p;
        if ((int)t.size() < 2 * p + 2) throw Error("knot vector too short for the degree");
        for (size_t i = 1; i < t.size(); ++i)
            if (t[i] < t[i - 1]) throw Error("knot vector must be nondecreasing");
        for (int i = 1; i <= p; ++i)
            if (t[i] != t[0] || t[t.size() - 1 - i] != t.back()) throw Error("knot vector must be open");
        for (size_t i = p + 1; i + p + 1 < t.size(); ++i) {
            size_t m = 1;
            while (i + m < t.size() - p - 1 && t[i + m] == t[i]) ++m;
            if ((int)m > p) throw Error("interior knot multiplicity above p (discontinuous basis)");
        }
        k0[pr.dim - 1] = t;
    }
    if (pr.nlayers > 0) {
        if (pr.dim != 3) throw Error("layers need dim = 3 (the layered axis is z)");
        if (!pr.layer_bounds || !pr.layer_mass || !pr.layer_v) throw Error("layer arrays missing");
        if (pr.cap_kind) throw Error("a CAP with layers is not supported");
        if (pr.pot_kind == 1) throw Error("the harmonic potential with layers is not supported");
        if (pr.dirichlet) throw Error("Dirichlet walls with layers are not supported");
        c->layered = true;
        c->lay_b.assign(pr.layer_bounds, pr.layer_bounds + pr.nlayers + 1);
        c->lay_m.assign(pr.layer_mass, pr.layer_mass + pr.nlayers);
        c->lay_v.assign(pr.layer_v, pr.layer_v + pr.nlayers);
        for (int i = 0; i < pr.nlayers; ++i) {
            if (!(c->lay_m[i] > 0)) throw Error("layer masses must be positive");
            if (!(c->lay_b[i + 1] > c->lay_b[i])) throw Error("layer bounds must increase");
        }
    }
    knots = level_knots(k0, pr.dim, pr.p, pr.mg_levels);
    const int nl = (int)knots.size();
    for (int l = 0; l < nl; ++l) {
        auto L = std::make_unique<Level>();
        for (int d = 0; d < 3; ++d) {
            if (d < pr.dim) tdz::build_axis(L->ax[d], knots[l][d], pr.p, pr.nquad, pr.dirichlet != 0, c->ph);
            else tdz::build_degenerate_axis(L->ax[d]);
            L->n[d] = L->ax[d].n;
            L->M[d] = L->ax[d].M; L->S[d] = L->ax[d].S; L->Xb[d] = L->ax[d].Xb; L->Wc[d] = L->ax[d].Wc;
            L->Vs[d] = L->ax[d].Vs;
        }
        L->N = (long)L->n[0] * L->n[1] * L->n[2];
        L->layered = c->layered;
        if (c->layered && l == 0) layer_bands(c, L->ax[2], L->Mw, L->Sw, L->Vl);
        bool uniform = true;
        for (int d = 0; d < pr.dim; ++d)
            for (int e = 0; e < L->ax[d].E; ++e) uniform = uniform && L->ax[d].off[e] == e;
        L->simple = pr.dim == 3 && uniform;
        L->simple2 = pr.dim == 2 && uniform;
        c->lv.push_back(std::move(L));
    }
    // ---- transfers and Galerkin bands (P^T A P per axis; exact for separable terms)
    for (int l = 0; l + 1 < nl; ++l) {
        Level &F = *c->lv[l], &Cl = *c->lv[l + 1];
        for (int d = 0; d < 3; ++d) {
            std::vector<std::vector<std::pair<int, double>>> P;
            if (d < pr.dim) {
                P = tdz::refinement_matrix(knots[l + 1][d], knots[l][d], pr.p);
                // the coarse B-splines must be reproduced by the fine ones through P (the
                // reference checks partition of unity and linear precision, src/multigrid.cpp:283-292)
                const double err = tdz::refinement_error(knots[l + 1][d], knots[l][d], pr.p, P);
                if (!(err <= 1e-11)) throw Error("knot-insertion prolongation is not exact");
                // walls: the coarse wall function maps only to the fine wall function
                if (pr.dirichlet) {
                    for (size_t i = 0; i < P.size(); ++i) {
                        auto &r = P[i];
                        r.erase(std::remove_if(r.begin(), r.end(), [&](auto &e) {
                                    return Cl.ax[d].isWall[e.first] || F.ax[d].isWall[i];
                                }), r.end());
                    }
                }
            } else {
                P = {{{0, 1.0}}};
            }
            const int nf = F.n[d], nc = Cl.n[d];
            if (d < pr.dim) {
                Cl.M[d] = galerkin_band(F.M[d], nf, pr.p, P, nc);
                Cl.S[d] = galerkin_band(F.S[d], nf, pr.p, P, nc);
                Cl.Xb[d] = galerkin_band(F.Xb[d], nf, pr.p, P, nc);
                Cl.Wc[d] = galerkin_band(F.Wc[d], nf, pr.p, P, nc);
                Cl.Vs[d] = galerkin_band(F.Vs[d], nf, pr.p, P, nc);
                if (c->layered && d == 2) {
                    Cl.Mw = galerkin_band(F.Mw, nf, pr.p, P, nc);
                    Cl.Sw = galerkin_band(F.Sw, nf, pr.p, P, nc);
                    Cl.Vl = galerkin_band(F.Vl, nf, pr.p, P, nc);
                }
                if (pr.dirichlet)
                    for (int i = 0; i < nc; ++i)
                        if (Cl.ax[d].isWall[i]) {
                            const int bw = 2 * pr.p + 1;
                            for (int k = 0; k < bw; ++k) {
                                Cl.M[d][(size_t)i * bw + k] = 0; Cl.S[d][(size_t)i * bw + k] = 0;
                                Cl.Xb[d][(size_t)i * bw + k] = 0; Cl.Wc[d][(size_t)i * bw + k] = 0;
                                Cl.Vs[d][(size_t)i * bw + k] = 0;
                            }
                        }
            }
            std::vector<std::vector<std::pair<int, double>>> R(nc);
            for (int i = 0; i < nf; ++i)
                for (auto &e : P[i]) R[e.first].push_back({i, e.second});
            F.P[d] = make_csr(c, P);
            F.R[d] = make_csr(c, R);
        }
    }
    for (auto &L : c->lv) upload_level_tables(c, *L);
    autotune_pot(c, *c->lv[0]);   // coarse levels apply the Galerkin stencil
    const long N = c->lv[0]->N;
    for (int i = 0; i < 4; ++i) c->sd.tmp[i] = track(c, dalloc<dcomplex>(N));
    c->sd.w1 = track(c, dalloc<dcomplex>(N));
    c->sd.w2 = track(c, dalloc<dcomplex>(N));
    c->rhs = track(c, dalloc<dcomplex>(N));
    c->restart = pr.restart > 0 ? pr.restart : (N > 100000 ? 50 : 30);
    if (c->restart > 60) throw Error("restart must be <= 60");
    c->Vb = track(c, dalloc<dcomplex>((size_t)(c->restart + 1) * N));
    c->Zb = track(c, dalloc<dcomplex>((size_t)c->restart * N));
    c->dpart = track(c, dalloc<dcomplex>((size_t)(c->restart + 1) * tdz::kRedBlocks));
    // dh, dh2 and dnorm in one block: an Arnoldi step reads them back with one copy
    c->dh = track(c, dalloc<dcomplex>(2 * (c->restart + 1) + 1));
    c->dh2 = c->dh + (c->restart + 1);
    c->dpartr = track(c, dalloc<double>(tdz::kRedBlocks));
    c->dnorm = reinterpret_cast<double *>(c->dh + 2 * (c->restart + 1));
    c->prob.knots = nullptr;   // the caller's arrays are not kept
    c->prob.layer_bounds = c->prob.layer_mass = c->prob.layer_v = nullptr;
    const double zero[3] = {0, 0, 0};
    set_shift(c, 1.0, 0.0, zero);
    // device data that later calls would otherwise allocate at first use: the potential
    // kernel's scratch, the complex position and zero bands of the dipole products, the zero
    // vector of the wall rows (tdz_observables_noalloc allocates nothing)
    if (c->ph.pot_kind) pot_scratch<double>(c, *c->lv[0]);
    auto &pd = c->pred;
    Level &L0 = *c->lv[0];
    for (int d = 0; d < 3; ++d) {
        if (!pd.Zband[d]) pd.Zband[d] = upload(c, std::vector<dcomplex>(L0.M[d].size(), dcomplex{0, 0}));
        std::vector<dcomplex> B(L0.Xb[d].size());
        for (size_t k = 0; k < B.size(); ++k) B[k] = {L0.Xb[d][k], 0.0};
        if (!pd.Xband[d]) pd.Xband[d] = upload(c, B);
    }
    if (pr.dirichlet && !c->zerov) c->zerov = upload(c, std::vector<dcomplex>(L0.N, dcomplex{0, 0}));
    tdz::noalloc_init();
}

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

const char *tdz_last_error(void) { return g_err.c_str(); }
const char *tdz_last_warning(void) { return g_warn.c_str(); }

int tdz_create(const tdz_problem *prob, tdz_ctx **ctx)
{
    if (!prob || !ctx) { g_err = "null argument"; return 1; }
    auto *c = new tdz_ctx();
    const int rc = guarded([&] { build(c, *prob); });
    if (rc) {
        tdz_destroy(c);
        *ctx = nullptr;
        return rc;
    }
    *ctx = c;
    return 0;
}

int tdz_destroy(tdz_ctx *c)
{
    if (!c) return 0;
    if (c->st) hipStreamSynchronize(c->st);
    for (auto &e : c->ev) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    for (void *p : c->allocs) hipFree(p);
    for (double *p : c->user_vecs) hipFree(p);
    if (c->hstage) hipHostFree(c->hstage);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
    return 0;
}

long tdz_size(const tdz_ctx *c) { return c ? c->lv[0]->N : 0; }
int tdz_dims(const tdz_ctx *c, int n[3])
{
    if (!c) return 1;
    for (int d = 0; d < 3; ++d) n[d] = c->lv[0]->n[d];
    return 0;
}
int tdz_levels(const tdz_ctx *c) { return c ? (int)c->lv.size() : 0; }

int tdz_vec_alloc(tdz_ctx *c, double **v)
{
    return guarded([&] {
        dcomplex *p = dalloc<dcomplex>(c->lv[0]->N);
        HIPCHK(hipMemset(p, 0, c->lv[0]->N * sizeof(dcomplex)));
        c->user_vecs.push_back(reinterpret_cast<double *>(p));
        *v = reinterpret_cast<double *>(p);
    });
}
int tdz_vec_free(tdz_ctx *c, double *v)
{
    return guarded([&] {
        auto it = std::find(c->user_vecs.begin(), c->user_vecs.end(), v);
        if (it == c->user_vecs.end()) throw Error("vector not owned by this context");
        c->user_vecs.erase(it);
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipFree(v));
    });
}
int tdz_vec_upload(tdz_ctx *c, double *d, const double *h)
{
    return guarded([&] {
        HIPCHK(hipMemcpyAsync(d, h, c->lv[0]->N * sizeof(dcomplex), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}
int tdz_vec_download(tdz_ctx *c, const double *d, double *h)
{
    return guarded([&] {
        HIPCHK(hipMemcpyAsync(h, d, c->lv[0]->N * sizeof(dcomplex), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}
int tdz_sync(tdz_ctx *c)
{
    return guarded([&] { HIPCHK(hipStreamSynchronize(c->st)); });
}

int tdz_set_shift(tdz_ctx *c, double ar, double ai, double gr, double gi, const double sE[3])
{
    return guarded([&] { set_shift(c, cplx(ar, ai), cplx(gr, gi), sE); });
}

int tdz_op_apply(tdz_ctx *c, const double *x, double *y)
{
    return guarded([&] {
        level_apply(c, *c->lv[0], reinterpret_cast<const dcomplex *>(x), reinterpret_cast<dcomplex *>(y));
        HIPCHK(hipGetLastError());
    });
}

int tdz_term_apply(tdz_ctx *c, int term, const double *x, double *y)
{
    return guarded([&] {
        if (term < TDZ_TERM_M || term > TDZ_TERM_DZ) throw Error("unknown term");
        const cplx a0 = c->alpha, g0 = c->g;
        double s0[3] = {c->sE[0], c->sE[1], c->sE[2]};
        tdz::PhysParams keep = c->ph;
        Level &L = *c->lv[0];
        const dcomplex *xd = reinterpret_cast<const dcomplex *>(x);
        dcomplex *yd = reinterpret_cast<dcomplex *>(y);
        // one term = a shift with only that term switched on
        const double inv2m = 0.5 / c->ph.mass;
        std::vector<std::vector<dcomplex>> B(3);
        for (int d = 0; d < 3; ++d) {
            const size_t sz = L.M[d].size();
            B[d].assign(sz, {0, 0});
            const bool act = d < c->prob.dim;
            for (size_t k = 0; k < sz; ++k) {
                cplx b = 0;
                if (L.layered) {   // x and y terms carry the 1/m z band (Mz2) except D_x, D_y
                    if (term == TDZ_TERM_M && d == 2) b = L.M[2][k];
                    if (term == TDZ_TERM_K) b = d < 2 ? 0.5 * L.S[d][k] : 0.5 * L.Sw[k];
                    if (term == TDZ_TERM_V && d == 2) b = c->ph.s * L.Vl[k];
                } else {
                    if (term == TDZ_TERM_M && d == 0) b = L.M[d][k];
                    if (term == TDZ_TERM_K && act) b = inv2m * L.S[d][k];
                    if (term == TDZ_TERM_CAP && act) b = cplx(0, -1) * L.Wc[d][k];
                }
                if (term >= TDZ_TERM_DX && d == term - TDZ_TERM_DX && act) b = L.Xb[d][k];
                B[d][k] = dc(b);
            }
            HIPCHK(hipMemcpyAsync(L.d.B[d], B[d].data(), sz * sizeof(dcomplex), hipMemcpyHostToDevice, c->st));
        }
        tdz::SepBandsT<double> sb = sep_bands<double>(L);
        if (term >= TDZ_TERM_DX) sb.Mz2 = nullptr;
        if (term == TDZ_TERM_V) {
            if (L.layered) tdz::launch_sep_apply<double>(c->st, sb, xd, yd, c->sd.tmp);   // s V_l(z)
            else HIPCHK(hipMemsetAsync(yd, 0, L.N * sizeof(dcomplex), c->st));
            if (c->ph.pot_kind) pot_apply(c, L, 1.0, xd, yd);
        } else {
            tdz::launch_sep_apply<double>(c->st, sb, xd, yd, c->sd.tmp);
        }
        if (c->prob.dirichlet) {
            const unsigned char *w[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
            tdz::launch_wall_fix<double>(c->st, L.n, w, xd, yd);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
        c->ph = keep;
        set_shift(c, a0, g0, s0);
    });
}

int tdz_pc_setup(tdz_ctx *c, int kind)
{
    return guarded([&] { pc_setup(c, kind); });
}

int tdz_pc_apply(tdz_ctx *c, const double *r, double *z)
{
    return guarded([&] {
        if (!c->pc_ready) throw Error("preconditioner not set up");
        pc_apply_dev(c, reinterpret_cast<const dcomplex *>(r), reinterpret_cast<dcomplex *>(z));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int tdz_ksp_solve(tdz_ctx *c, const double *b, double *x, double rtol, int maxit, tdz_ksp_stats *stats)
{
    return guarded([&] {
        if (!c->pc_ready) pc_setup(c, c->pc_kind);
        fgmres(c, reinterpret_cast<const dcomplex *>(b), reinterpret_cast<dcomplex *>(x), rtol, maxit, stats);
        HIPCHK(hipStreamSynchronize(c->st));   // the results are final when the call returns
    });
}

int tdz_ifunction(tdz_ctx *c, const double sE[3], const double *psi, const double *psidot, double *F)
{
    return guarded([&] {
        if (!psi || !psidot || !F) throw Error("null vector");
        Level &L = *c->lv[0];
        const long N = L.N;
        const cplx a0 = c->alpha, g0 = c->g;
        const double s0[3] = {c->sE[0], c->sE[1], c->sE[2]};
        const dcomplex *x = reinterpret_cast<const dcomplex *>(psi), *xd = reinterpret_cast<const dcomplex *>(psidot);
        dcomplex *y = reinterpret_cast<dcomplex *>(F);
        set_shift(c, 0.0, -1.0, sE);   // F = -(H + sE.D) psi
        tdz::launch_sep_apply<double>(c->st, sep_bands<double>(L), x, y, c->sd.tmp);
        if (c->ph.pot_kind) pot_apply<double>(c, L, -1.0, x, y);
        mass_apply(c, xd, c->sd.w1);    // + i M psidot
        tdz::launch_axpby(c->st, {0, 1}, c->sd.w1, {1, 0}, y, N);
        if (c->prob.dirichlet) {         // wall DOFs are not unknowns of the propagation: F = 0 there
            if (!c->zerov) c->zerov = upload(c, std::vector<dcomplex>(N, dcomplex{0, 0}));
            const unsigned char *w[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
            tdz::launch_wall_fix<double>(c->st, L.n, w, c->zerov, y);
        }
        HIPCHK(hipGetLastError());
        set_shift(c, a0, g0, s0);
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int tdz_level_dims(const tdz_ctx *c, int level, int n[3])
{
    if (!c || !n || level < 0 || level >= (int)c->lv.size()) { g_err = "no such level"; return 1; }
    for (int d = 0; d < 3; ++d) n[d] = c->lv[level]->n[d];
    return 0;
}

int tdz_mg_transfer(tdz_ctx *c, int level, int dir, const double *in, double *out)
{
    return guarded([&] {
        if (level < 0 || level + 1 >= (int)c->lv.size()) throw Error("no coarser level below this one");
        if (dir != 0 && dir != 1) throw Error("dir must be 0 (restrict) or 1 (prolong)");
        Level &F = *c->lv[level], &Cl = *c->lv[level + 1];
        const dcomplex *a = reinterpret_cast<const dcomplex *>(in);
        dcomplex *b = reinterpret_cast<dcomplex *>(out);
        if (dir == 0) transfer<double>(c, F.R, F.n, Cl.n, a, b, c->sd.w1, c->sd.w2, false);
        else transfer<double>(c, F.P, Cl.n, F.n, a, b, c->sd.w1, c->sd.w2, false);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int tdz_level_apply(tdz_ctx *c, int level, const double *x, double *y)
{
    return guarded([&] {
        if (level < 0 || level >= (int)c->lv.size()) throw Error("no such level");
        refresh_coarse(c);
        level_apply<double>(c, *c->lv[level], reinterpret_cast<const dcomplex *>(x), reinterpret_cast<dcomplex *>(y));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int tdz_level_solve(tdz_ctx *c, int level, const double *b, double *x)
{
    return guarded([&] {
        const int last = (int)c->lv.size() - 1;
        if (level != last || level < 1) throw Error("the level solve takes the coarsest level of a hierarchy");
        Level &L = *c->lv[level];
        ensure_galerkin(c);
        setup_fd(c, L);
        refresh_coarse(c);
        vcycle<double>(c, level, reinterpret_cast<const dcomplex *>(b), reinterpret_cast<dcomplex *>(x));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int tdz_pc_set_corrections(tdz_ctx *c, int sweeps)
{
    return guarded([&] {
        if (sweeps < 1 || sweeps > 4) throw Error("correction sweeps must be 1..4");
        c->corrections = sweeps;
    });
}

int tdz_mg_set_smoother(tdz_ctx *c, int kind, int degree)
{
    return guarded([&] {
        if (kind != 0 && kind != 1) throw Error("smoother kind must be 0 (separable solve) or 1 (Jacobi-Chebyshev)");
        if (degree < 1 || degree > 8) throw Error("Chebyshev degree must be 1..8");
        c->smoother = kind;
        c->cheb_degree = degree;
        if (kind == 1 && c->pc_ready &&
            (c->pc_kind == TDZ_PC_MG || c->pc_kind == TDZ_PC_MG_MIXED || c->pc_kind == TDZ_PC_LEARNED))
            cheb_setup(c);
    });
}

int tdz_set_fd64_transform(tdz_ctx *c, int kind)
{
    return guarded([&] {
        if (kind != 0 && kind != 1) throw Error("FP64 transform kind must be 0 (VALU) or 1 (FP64 MFMA)");
        c->fd64_mfma = kind == 1;
    });
}

int tdz_set_fd_transform(tdz_ctx *c, int kind)
{
    return guarded([&] {
        if (kind < 0 || kind > 3)
            throw Error("transform kind must be 0 (FP32 MFMA), 1 (split FP16 MFMA), 2 (split FP16 MFMA, matrix in "
                        "LDS) or 3 (single FP16 products, for corrections)");
        c->fd_transform = kind;
    });
}

int tdz_set_correction_form(tdz_ctx *c, int form)
{
    return guarded([&] {
        if (form != 0 && form != 1) throw Error("correction form must be 0 (whole iterate) or 1 (correction alone)");
        c->corr_form = form;
    });
}

int tdz_mg_set_coarse(tdz_ctx *c, int kind, double rtol, int maxit)
{
    return guarded([&] {
        if (kind != 0 && kind != 1) throw Error("coarse solve kind must be 0 (separable sweeps) or 1 (Krylov)");
        if (kind == 1 && !(rtol > 0 && rtol < 1)) throw Error("coarse solve rtol must be in (0, 1)");
        if (kind == 1 && (maxit < 1 || maxit > 4096)) throw Error("coarse solve maxit must be 1..4096");
        c->coarse_kind = kind;
        if (kind == 1) {
            c->coarse_rtol = rtol;
            c->coarse_maxit = maxit;
        }
    });
}

int tdz_mg_coarse_status(const tdz_ctx *c, int *its, double *res)
{
    return guarded([&] {
        if (its) *its = c->ck.last_its;
        if (res) *res = c->ck.last_res;
    });
}

int tdz_set_pc_refresh(tdz_ctx *c, int max_its)
{
    return guarded([&] {
        c->refresh_cap = max_its < 0 ? -1 : max_its;
        c->refresh_base = -1;
    });
}

int tdz_coarsen_knots(const double *U, int m, double *C, int *mc)
{
    return guarded([&] {
        if (!U || !C || !mc || m < 2) throw Error("bad knot vector");
        std::vector<double> c;
        const bool removed = tdz::coarsen_knots(std::vector<double>(U, U + m), c);
        *mc = removed ? (int)c.size() : 0;
        if (removed) std::copy(c.begin(), c.end(), C);
    });
}

int tdz_refinement_matrix(const double *C, int mc, const double *F, int mf, int p, double *R)
{
    return guarded([&] {
        if (!C || !F || !R || p < 1 || mc < 2 * p + 2 || mf < mc) throw Error("bad knot vectors");
        const auto rows = tdz::refinement_matrix(std::vector<double>(C, C + mc), std::vector<double>(F, F + mf), p);
        const int nf = mf - p - 1, nc = mc - p - 1;
        std::fill(R, R + (size_t)nf * nc, 0.0);
        for (int i = 0; i < nf; ++i)
            for (auto &e : rows[i]) R[(size_t)i * nc + e.first] = e.second;
    });
}

}  // extern "C"

namespace {
// Checks of a multigrid run, each warned once per context (stderr and tdz_last_warning).
// Strong field (TDSEZSetupMultigrid, src/multigrid.cpp:213-237): once tau max|E.r| exceeds
// 1.5 over the pulse the field term dominates M + i tau H_E and the coarse levels no longer
// represent it. tau = dt/2 (CN), dt/sqrt(12) (Pade(2,2)), dt/(2 sqrt(12)) (Magnus 4, half
// steps). Dominant CAP (this build's smoother): the V-cycle smooths with the separable
// solve of M + g K, which leaves the CAP out; where tau max W exceeds 1/2 (a profile that
// diverges at the walls, as the Manolopoulos one does) the smoothing sweep amplifies the
// error there and the V-cycle may stall.
void multigrid_checks(tdz_ctx *c, const tdz_pulse *pulse, double tau, double dt)
{
    const bool mg = c->pc_kind == TDZ_PC_MG || c->pc_kind == TDZ_PC_MG_MIXED ||
                    (c->pc_kind == TDZ_PC_LEARNED && c->learn.fallback_kind == TDZ_PC_MG);
    if (!mg || dt <= 0) return;
    auto warn = [&](const char *msg) {
        g_warn = msg;
        std::fprintf(stderr, "tdz: warning: %s\n", msg);
    };
    char msg[240];
    if (!c->cap_warned && c->cap_on) {
        double w = 0;
        for (int d = 0; d < c->prob.dim; ++d)
            for (double v : c->lv[0]->ax[d].cap) w = std::max(w, std::fabs(v));
        if (tau * w > 0.5) {
            std::snprintf(msg, sizeof msg,
                          "tau*max W = %.2g > 0.5: the CAP dominates the propagator operator near the walls and "
                          "the multigrid smoother (separable solve without the CAP) may not converge (use the "
                          "separable preconditioner)", tau * w);
            warn(msg);
            c->cap_warned = true;
        }
    }
    if (!pulse || c->field_warned || pulse->omega <= 0 || pulse->ncycles <= 0) return;
    // the scan below samples the whole pulse: once per (pulse, tau, dt)
    const double key[8] = {pulse->E0[0], pulse->E0[1], pulse->E0[2], pulse->omega, pulse->ncycles, pulse->cep, tau, dt};
    if (std::equal(key, key + 8, c->field_checked)) return;
    std::copy(key, key + 8, c->field_checked);
    const double T = pulse->ncycles * 2.0 * M_PI / pulse->omega;
    double Lmax[3] = {0, 0, 0};
    for (int a = 0; a < c->prob.dim; ++a) Lmax[a] = std::max(std::fabs(c->prob.lo[a]), std::fabs(c->prob.hi[a]));
    const long n = (long)std::ceil(4.0 * T / dt);
    double worst = 0;
    for (long i = 0; i <= n; ++i) {
        const double t = T * (double)i / (double)std::max(n, 1L);
        double v = 0;
        for (int a = 0; a < c->prob.dim; ++a) v += std::fabs(pulse_field(*pulse, a, t)) * Lmax[a];
        worst = std::max(worst, v);
    }
    if (tau * worst > 1.5) {
        std::snprintf(msg, sizeof msg,
                      "tau*max|E.r| = %.2g > 1.5: the field term dominates the propagator operator and multigrid "
                      "may not converge (use the separable preconditioner or a smaller time step)",
                      tau * worst);
        warn(msg);
        c->field_warned = true;
    }
}

// Preconditioner refresh rule of TDSEZPCRefresh (include/tdsez_propagator.hpp:17-31): the
// first solve after a build sets base; a solve above 2 base + 5 iterations, or above the
// cap (-tdsez_pc_refresh_its) when one is set, asks for a rebuild before the next solve.
bool refresh_due(tdz_ctx *c, int &base, int its)
{
    const bool over_cap = c->refresh_cap >= 0 && its > c->refresh_cap;
    if (!over_cap) {
        if (base < 0) {
            base = its;
            return false;
        }
        if (its <= 2 * base + 5) return false;
    }
    base = -1;
    ++c->rebuilds;
    return true;
}

// Time-step solve with the current preconditioner. The learned preconditioner has
// regressed when a solve needs more iterations than the same preconditioner without
// the network took on the training steps: such a solve is stopped at that count, plain
// multigrid takes over for the rest of the run and finishes it from the last iterate,
// so the answer is always judged by the same FP64 residual.
void step_solve(tdz_ctx *c, const dcomplex *b, dcomplex *x, double rtol, int maxit, tdz_ksp_stats *st,
                bool r0_in_v0, dcomplex *ax_out = nullptr)
{
    LearnState &ls = c->learn;
    if (c->pc_kind != TDZ_PC_LEARNED || ls.fallen || ls.base_its < 0) {
        fgmres(c, b, x, rtol, maxit, st, r0_in_v0, false, ax_out);
        return;
    }
    tdz_ksp_stats s1{};
    if (fgmres(c, b, x, rtol, std::min(maxit, ls.base_its), &s1, r0_in_v0, true, ax_out) >= 0) {
        *st = s1;
        return;
    }
    ls.fallen = true;
    ++ls.fallbacks;
    fgmres(c, b, x, rtol, std::max(maxit - s1.its, 1), st, false, false, ax_out);
    st->its += s1.its;
    st->restarts += s1.restarts;
}

// One CN step psi <- (M + g H_E)^-1 (M - g H_E) psi, g = i dt/2 (src/propagator.cpp:470-540),
// on its own (tdz_step_cn): the right-hand side is an operator apply. The steps of an evolve
// run go through cn_step_pred, which forms it from the products its last solve left.
void cn_step(tdz_ctx *c, dcomplex *x, double dt, const double sE_mid[3], double rtol, int maxit, tdz_ksp_stats *st)
{
    if (!c->pc_ready) pc_setup(c, c->pc_kind);
    Level &L = *c->lv[0];
    const long N = L.N;
    const cplx g(0.0, 0.5 * dt);
    c->pred.count = 0;   // the state moves on without the predictor's bookkeeping
    tdz::SepBandsT<double> mb = sep_bands<double>(L);
    for (int d = 0; d < 3; ++d) mb.B[d] = L.d.Bmass[d];
    const unsigned char *wall[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
    set_shift(c, 1.0, -g, sE_mid);        // rhs = (M - i dt/2 H_E) psi
    level_apply(c, L, x, c->rhs);
    tdz::launch_sep_apply<double>(c->st, mb, x, c->sd.w1, c->sd.tmp);
    if (c->prob.dirichlet) tdz::launch_wall_fix<double>(c->st, L.n, wall, x, c->sd.w1);
    // (M + i dt/2 H_E) psi = 2 M psi - rhs, so the initial residual of the guess psi
    // is r0 = 2 (rhs - M psi): one mass apply instead of a full operator apply
    // (the same identity as the stage right-hand sides of src/propagator.cpp:589-624)
    HIPCHK(hipMemcpyAsync(c->Vb, c->rhs, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
    tdz::launch_axpby(c->st, {-2.0, 0}, c->sd.w1, {2.0, 0}, c->Vb, N);
    set_shift(c, 1.0, g, sE_mid);         // solve (M + i dt/2 H_E) psi' = rhs, guess psi
    step_solve(c, c->rhs, x, rtol, maxit, st, true);
}

// ------------------------------------------------------------------ CN predictor
// y = D_a x, the dipole along axis a: the position band on axis a and the mass bands on
// the others; zero on Dirichlet walls, whose rows of the propagator matrices carry no field.
void dipole_apply(tdz_ctx *c, int a, const dcomplex *x, dcomplex *y)
{
    Level &L = *c->lv[0];
    auto &p = c->pred;
    for (int d = 0; d < 3; ++d)
        if (!p.Zband[d]) p.Zband[d] = upload(c, std::vector<dcomplex>(L.M[d].size(), dcomplex{0, 0}));
    if (!p.Xband[a]) {
        std::vector<dcomplex> B(L.Xb[a].size());
        for (size_t k = 0; k < B.size(); ++k) B[k] = {L.Xb[a][k], 0.0};
        p.Xband[a] = upload(c, B);
    }
    tdz::SepBandsT<double> db = sep_bands<double>(L);
    db.Mz2 = nullptr;   // layered structures: the dipole carries the plain mass along z
    for (int d = 0; d < 3; ++d) db.B[d] = d == a ? p.Xband[a] : p.Zband[d];
    tdz::launch_sep_apply<double>(c->st, db, x, y, c->sd.tmp);
    if (c->prob.dirichlet) {
        if (!c->zerov) c->zerov = upload(c, std::vector<dcomplex>(L.N, dcomplex{0, 0}));
        const unsigned char *wall[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
        tdz::launch_wall_fix<double>(c->st, L.n, wall, c->zerov, y);
    }
}

// Buffers of history slot h.
void pred_alloc(tdz_ctx *c, int h)
{
    auto &p = c->pred;
    const long N = c->lv[0]->N;
    if (!p.psi[h]) {
        p.psi[h] = track(c, dalloc<dcomplex>(N));
        p.P[h] = track(c, dalloc<dcomplex>(N));
    }
    if (!p.Mx[h]) p.Mx[h] = track(c, dalloc<dcomplex>(N));
    for (int a = 0; a < 3; ++a)
        if ((p.axes >> a & 1) && !p.D[h][a]) p.D[h][a] = track(c, dalloc<dcomplex>(N));
}

// Push state x into the predictor history (slot (head + 1) % ring) with M x, D_a x and
// P = (M + g H_sE) x: with ax_in_slot the solve that produced x already left its A x in
// that slot's P (under the field sE of its step), else P is one operator apply without
// field.
void pred_push(tdz_ctx *c, const dcomplex *x, bool ax_in_slot, const double sE[3])
{
    auto &p = c->pred;
    Level &L = *c->lv[0];
    const long N = L.N;
    const cplx g(0.0, 0.5 * p.dt);
    const int h = (p.head + 1) % p.ring;
    pred_alloc(c, h);
    // x, M x and D_a x: one fused pass for a field along z alone, else a copy and one apply each
    bool fused = false;
    if (p.axes == 4u && c->prob.dim == 3 && L.ax[0].p == L.ax[2].p && L.ax[1].p == L.ax[2].p) {
        if (!p.Xz) p.Xz = upload(c, L.Xb[2]);
        const double *M[3] = {L.d.M[0], L.d.M[1], L.d.M[2]};
        fused = tdz::launch_mass_dipole_z(c->st, L.n, L.ax[2].p, M, p.Xz, x, p.Mx[h], p.D[h][2], p.psi[h]);
        if (fused && c->prob.dirichlet) {   // walls: identity rows of M, no field
            if (!c->zerov) c->zerov = upload(c, std::vector<dcomplex>(L.N, dcomplex{0, 0}));
            const unsigned char *wall[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
            tdz::launch_wall_fix<double>(c->st, L.n, wall, c->zerov, p.D[h][2]);
        }
    }
    if (!fused) {
        HIPCHK(hipMemcpyAsync(p.psi[h], x, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
        for (int a = 0; a < 3; ++a)
            if (p.axes >> a & 1) dipole_apply(c, a, x, p.D[h][a]);
        tdz::SepBandsT<double> mb = sep_bands<double>(L);
        for (int d = 0; d < 3; ++d) mb.B[d] = L.d.Bmass[d];
        tdz::launch_sep_apply<double>(c->st, mb, x, p.Mx[h], c->sd.tmp);
    }
    if (c->prob.dirichlet) {
        const unsigned char *wall[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
        tdz::launch_wall_fix<double>(c->st, L.n, wall, x, p.Mx[h]);
    }
    if (ax_in_slot) {
        for (int a = 0; a < 3; ++a) p.sEh[h][a] = sE[a];
    } else {
        const double zero[3] = {0, 0, 0};
        set_shift(c, 1.0, g, zero);
        level_apply(c, L, x, p.P[h]);
        for (int a = 0; a < 3; ++a) p.sEh[h][a] = 0.0;
        p.since_sync = 0;
    }
    p.head = h;
    p.count = std::min(p.count + 1, p.ring);
}

// True when x is bit for bit the newest state of the history (an evolve call that
// continues the run of the previous one).
bool pred_continues(tdz_ctx *c, const dcomplex *x, double dt)
{
    auto &p = c->pred;
    if (p.count == 0 || p.dt != dt) return false;
    const long N = c->lv[0]->N;
    tdz::LinComb d;
    d.add({1, 0}, x);
    d.add({-1, 0}, p.psi[p.head]);
    tdz::launch_lincomb(c->st, d, c->sd.w2, N);
    tdz::launch_norm2(c->st, c->sd.w2, N, c->dpartr, c->dnorm);
    double v = -1;
    HIPCHK(hipMemcpyAsync(&v, c->dnorm, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return v == 0.0;
}

// The stored products P_j are carried from solve to solve: each solve's A x comes from
// its Arnoldi relation and from the history products it started with, so their rounding
// errors propagate through the extrapolation. With weights of order q the recurrence
// delta_new = sum_j w_j delta_j has the characteristic polynomial (z - 1)^(q+1), and the
// inconsistency, injected at eps ||b|| per step, grows like n^(q+1/2) eps ||b|| (the
// solutions then drift from the true CN steps by about as much per step). Every
// sync_every steps the products are recomputed exactly (one operator apply per slot);
// the drift measured there steers the interval so that it stays between 1e-3 and 1e-2
// of the solve tolerance (a per-step state error well inside the tolerance's own).
void pred_resync(tdz_ctx *c, double rtol)
{
    auto &p = c->pred;
    if (p.order == 0 || p.count < 2 || ++p.since_sync < p.sync_every) return;
    p.since_sync = 0;
    Level &L = *c->lv[0];
    const long N = L.N;
    const cplx g(0.0, 0.5 * p.dt);
    double drift = 0;
    for (int j = 0; j < p.count; ++j) {
        const int s = (p.head - j + p.ring) % p.ring;
        set_shift(c, 1.0, g, p.sEh[s]);
        level_apply(c, L, p.psi[s], c->sd.w2);
        tdz::LinComb d;
        d.add({1, 0}, p.P[s]);
        d.add({-1, 0}, c->sd.w2);
        tdz::launch_lincomb(c->st, d, c->sd.w1, N);
        const double e2 = dev_norm2(c, c->sd.w1, N), p2 = dev_norm2(c, c->sd.w2, N);
        drift = std::max(drift, std::sqrt(e2 / std::max(p2, 1e-300)));
        HIPCHK(hipMemcpyAsync(p.P[s], c->sd.w2, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
    }
    if (drift < 1e-3 * rtol) p.sync_every = std::min(2 * p.sync_every, 512);
    else if (drift > 1e-2 * rtol) p.sync_every = std::max(p.sync_every / 2, 2);
    p.last_drift = drift;
}

// One CN step with an extrapolated initial guess. With the states psi_j (j = 0 the
// current one) of the history, the guess is the binomial extrapolation
// x0 = sum_j w_j psi_j, w = (1), (2, -1) or (3, -3, 1) by order; its residual
//   r0 = b - A x0,  A = M + g H_sE,  b = (M - g H_sE) psi_0 = 2 M psi_0 - A psi_0,
// follows from A psi_j = P_j + g sum_a (sE_a - sE^j_a) D_a psi_j (P_j = (M + g H_sE^j) psi_j,
// the A x its own solve returned) without an operator apply, so the step costs one mass
// apply and one dipole apply per field axis besides the solve.
// The solve is the same (M + g H_sE) psi' = b to the same tolerance as without the
// guess (src/propagator.cpp:78-251 starts from psi_0; the stage solves of
// src/propagator.cpp:629-638 extrapolate the same way).
void cn_step_pred(tdz_ctx *c, dcomplex *x, double dt, const double sE[3], double rtol, int maxit, tdz_ksp_stats *st)
{
    if (!c->pc_ready) pc_setup(c, c->pc_kind);
    auto &p = c->pred;
    Level &L = *c->lv[0];
    const long N = L.N;
    const cplx g(0.0, 0.5 * dt);
    for (int a = 0; a < c->prob.dim; ++a)
        if (sE[a] != 0 && !(p.axes >> a & 1)) {   // a new field axis: restart the history
            p.axes |= 1u << a;
            p.count = 0;
        }
    if (p.count == 0 || p.dt != dt) {
        p.count = 0;
        p.dt = dt;
        pred_push(c, x, false, sE);
    }
    pred_resync(c, rtol);
    auto slot = [&](int j) { return (p.head - j + p.ring) % p.ring; };
    const int h0 = slot(0);
    tdz::LinComb rb;   // b = 2 M psi_0 - P_0 - g (sE - sE^0).D psi_0 (M psi_0 kept at its push)
    rb.add({2, 0}, p.Mx[h0]);
    rb.add({-1, 0}, p.P[h0]);
    for (int a = 0; a < 3; ++a)
        if (p.axes >> a & 1) rb.add(dc(-g * (sE[a] - p.sEh[h0][a])), p.D[h0][a]);
    // binomial extrapolation weights of orders 0..4 (newest state first)
    static const double W[5][5] = {{1, 0, 0, 0, 0}, {2, -1, 0, 0, 0}, {3, -3, 1, 0, 0}, {4, -6, 4, -1, 0},
                                   {5, -10, 10, -5, 1}};
    const int o = std::min(p.order, p.count - 1);
    tdz::LinComb r0, x0;   // r0 = b - sum_j w_j A psi_j;  x0 = sum_j w_j psi_j
    for (int j = 0; j <= o; ++j) {
        const int s = slot(j);
        const double w = W[o][j];
        r0.add({-w, 0}, p.P[s]);
        for (int a = 0; a < 3; ++a)
            if (p.axes >> a & 1) r0.add(dc(-w * g * (sE[a] - p.sEh[s][a])), p.D[s][a]);
        x0.add({w, 0}, p.psi[s]);
    }
    set_shift(c, 1.0, g, sE);
    // b, the solve's first basis vector before scaling, the guess and the norms of the
    // first two in one pass (dpart is free between solves: 2 (restart + 1) kRedBlocks
    // doubles); the solve that follows consumes the queued norms
    tdz::launch_lincomb3_norms(c->st, rb, r0, x0, c->rhs, c->Vb, x, N, reinterpret_cast<double *>(c->dpart),
                               2 * tdz::kRedBlocks, c->dnorm);
    c->norms_queued = true;
    // the solve leaves A x in the P of the slot the new state goes to (the oldest slot,
    // whose P the pass above has already read)
    const int hnext = (p.head + 1) % p.ring;
    pred_alloc(c, hnext);
    step_solve(c, c->rhs, x, rtol, maxit, st, true, p.P[hnext]);
    pred_push(c, x, true, sE);
}
}  // namespace

extern "C" {

int tdz_set_cn_predictor(tdz_ctx *c, int order)
{
    return guarded([&] {
        if (order < 0 || order > 4) throw Error("CN predictor order must be 0..4");
        c->pred.order = order;
        c->pred.ring = std::max(order + 1, 1);
        c->pred.count = 0;
        c->pred.head = 0;
    });
}

int tdz_cn_predictor_status(const tdz_ctx *c, int *sync_every, double *last_drift)
{
    if (!c) { g_err = "null context"; return 1; }
    if (sync_every) *sync_every = c->pred.sync_every;
    if (last_drift) *last_drift = c->pred.last_drift;
    return 0;
}

int tdz_step_cn(tdz_ctx *c, double *psi, double dt, const double sE_mid[3], double rtol, int maxit,
                tdz_ksp_stats *stats)
{
    return guarded([&] {
        tdz_ksp_stats st{};
        cn_step(c, reinterpret_cast<dcomplex *>(psi), dt, sE_mid, rtol, maxit, &st);
        if (stats) *stats = st;
        HIPCHK(hipStreamSynchronize(c->st));   // the results are final when the call returns
    });
}

}  // extern "C"

namespace {

// guess history buffers for the current order: 4 slots (2 stages x 2 Magnus halves)
void pade_hist_alloc(tdz_ctx *c)
{
    const long N = c->lv[0]->N;
    if (c->hist && c->hist_order == c->guess_order) return;
    if (c->hist) untrack(c, c->hist);
    c->hist = track(c, dalloc<dcomplex>((size_t)4 * c->guess_order * N));
    c->hist_order = c->guess_order;
    for (int s = 0; s < 4; ++s) c->hist_count[s] = 0;
    if (!c->xin) c->xin = track(c, dalloc<dcomplex>(N));
}

void pade22_step(tdz_ctx *c, dcomplex *x, double h, const double sE[3], int slot, double rtol, int maxit,
                 tdz_ksp_stats *stats)
{
    if (!c->pc_ready) pc_setup(c, c->pc_kind);
    Level &L = *c->lv[0];
    const long N = L.N;
    pade_hist_alloc(c);
    const int order = c->guess_order;
    const double s3 = std::sqrt(3.0);
    const cplx gk[2] = {h * cplx(-s3, 3.0) / 12.0, h * cplx(s3, 3.0) / 12.0};
    // x + sum_j w_j d_j, d_0 the newest increment: the order-q binomial extrapolation
    auto binom = [](int q, int j) {
        double b = 1;
        for (int i = 0; i < j + 1; ++i) b = b * (q - i) / (i + 1);
        return (j % 2 ? -1.0 : 1.0) * b;
    };
    tdz_ksp_stats tot{};
    for (int k = 0; k < 2; ++k) {
        const int sl = 2 * slot + k;
        dcomplex *H = c->hist + (size_t)sl * order * N;
        set_shift(c, 1.0, std::conj(gk[k]), sE);
        level_apply(c, L, x, c->rhs);
        HIPCHK(hipMemcpyAsync(c->xin, x, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
        const int q = c->guess_extrap ? std::min(order, c->hist_count[sl]) : 0;
        for (int j = 0; j < q; ++j) tdz::launch_axpby(c->st, {binom(q, j), 0}, H + (size_t)j * N, {1, 0}, x, N);
        set_shift(c, 1.0, gk[k], sE);
        tdz_ksp_stats st{};
        step_solve(c, c->rhs, x, rtol, maxit, &st, false);
        if (c->guess_extrap) {   // history: shift, newest increment y - x first
            for (int j = order - 1; j > 0; --j)
                HIPCHK(hipMemcpyAsync(H + (size_t)j * N, H + (size_t)(j - 1) * N, N * sizeof(dcomplex),
                                      hipMemcpyDeviceToDevice, c->st));
            HIPCHK(hipMemcpyAsync(H, x, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
            tdz::launch_axpby(c->st, {-1, 0}, c->xin, {1, 0}, H, N);
            c->hist_count[sl] = std::min(c->hist_count[sl] + 1, order);
        }
        // each stage keeps its own refresh tracker (src/propagator.cpp:650-668)
        if (refresh_due(c, c->stage_base[k], st.its)) pc_setup(c, c->pc_kind);
        tot.its += st.its;
        tot.restarts += st.restarts;
        tot.res = std::max(tot.res, st.res);
        tot.bnorm = st.bnorm;
    }
    tot.pc_rebuilds = c->rebuilds;
    if (stats) *stats = tot;
}

// L2 projection of a Gaussian packet (src/core.cpp:1188-1262): b_i = int B_i psi0,
// c = M^-1 b, c /= sqrt(c^H M c). psi0 is separable, so b is a Kronecker product
// of 1D integrals and M^-1 is applied exactly by fast diagonalisation with g = 0.
void project_gaussian(tdz_ctx *c, const double centre[3], const double width[3], const double mom[3],
                      dcomplex *out)
{
    Level &L = *c->lv[0];
    std::vector<cplx> bd[3];
    for (int d = 0; d < 3; ++d) {
        const tdz::Axis1D &A = L.ax[d];
        bd[d].assign(A.n, 0.0);
        if (d >= c->prob.dim) { bd[d][0] = 1.0; continue; }
        const int nen = A.p + 1;
        for (int e = 0; e < A.E; ++e)
            for (int q = 0; q < A.Q; ++q) {
                const double x = A.X[(size_t)e * A.Q + q], w = A.W[(size_t)e * A.Q + q];
                const double s = width[d], u = x - centre[d];
                const cplx v = std::pow(2.0 * M_PI * s * s, -0.25) *
                               std::exp(cplx(-u * u / (4.0 * s * s), mom[d] * x));
                for (int a = 0; a < nen; ++a)
                    bd[d][A.off[e] + a] += w * A.B0[((size_t)e * nen + a) * A.Q + q] * v;
            }
    }
    std::vector<dcomplex> b(L.N);
    const int n0 = L.n[0], n1 = L.n[1];
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i)
                b[i + (size_t)n0 * (j + (size_t)n1 * k)] = dc(bd[0][i] * bd[1][j] * bd[2][k]);
    HIPCHK(hipMemcpyAsync(c->sd.w1, b.data(), L.N * sizeof(dcomplex), hipMemcpyHostToDevice, c->st));
    setup_fd(c, L);
    const cplx a0 = c->alpha, g0 = c->g;
    double s0[3] = {c->sE[0], c->sE[1], c->sE[2]};
    c->alpha = 1.0;
    c->g = 0.0;
    fd_apply(c, L, c->sd.w1, out, c->sd.tmp[0], c->sd.tmp[1]);
    c->alpha = a0;
    c->g = g0;
    (void)s0;
    mass_apply(c, out, c->sd.w2);
    const double nrm = std::sqrt(dev_dot(c, out, c->sd.w2, L.N).real());
    tdz::launch_axpby(c->st, {1.0 / nrm, 0}, out, {0, 0}, out, L.N);
    HIPCHK(hipStreamSynchronize(c->st));
}

// Lowest eigenstate of H = K + sV (no CAP) by shifted inverse iteration with
// Rayleigh-quotient shift updates: (H - sigma M) x_{k+1} = M x_k by FGMRES with the
// separable preconditioner of (-sigma) M + (K + V_sep). Stops when the relative
// eigen-residual ||H x - E M x|| / ||M x|| < tol. Replaces the SLEPc shift-invert
// solve (src/core.cpp:758-1083) for the ground state only.
void ground_state(tdz_ctx *c, dcomplex *x, double sigma, double tol, int maxit, double *E_out, double *res_out,
                  int *its_out)
{
    Level &L = *c->lv[0];
    const long N = L.N;
    const int pc0 = c->pc_kind;
    const cplx a0 = c->alpha, g0 = c->g;
    double s0[3] = {c->sE[0], c->sE[1], c->sE[2]};
    const double zero[3] = {0, 0, 0};
    c->cap_on = false;
    c->pc_kind = TDZ_PC_SEPARABLE;
    setup_fd(c, L);
    auto rayleigh = [&](double &E, double &res) {
        mass_apply(c, x, c->sd.w2);                       // w2 = M x
        set_shift(c, 0.0, 1.0, zero);
        level_apply(c, L, x, c->sd.w1);                   // w1 = H x
        const double xm = dev_dot(c, x, c->sd.w2, N).real();
        E = dev_dot(c, x, c->sd.w1, N).real() / xm;
        tdz::launch_axpby(c->st, {-E, 0}, c->sd.w2, {1, 0}, c->sd.w1, N);   // w1 = H x - E M x
        res = std::sqrt(dev_norm2(c, c->sd.w1, N) / dev_norm2(c, c->sd.w2, N));
    };
    double E = 0, res = 1;
    int it = 0;
    try {
        rayleigh(E, res);
        for (; it < maxit && res > tol; ++it) {
            const double shift = it < 3 ? sigma : std::min(sigma, E - 10.0 * res);
            mass_apply(c, x, c->rhs);                  // rhs = M x
            set_shift(c, -shift, 1.0, zero);
            tdz_ksp_stats st{};
            fgmres(c, c->rhs, x, 1e-11, 500, &st);
            mass_apply(c, x, c->sd.w2);
            const double nrm = std::sqrt(dev_dot(c, x, c->sd.w2, N).real());
            tdz::launch_axpby(c->st, {1.0 / nrm, 0}, x, {0, 0}, x, N);
            rayleigh(E, res);
        }
    } catch (...) {
        c->cap_on = true;
        c->pc_kind = pc0;
        set_shift(c, a0, g0, s0);
        throw;
    }
    c->cap_on = true;
    c->pc_kind = pc0;
    set_shift(c, a0, g0, s0);
    if (E_out) *E_out = E;
    if (res_out) *res_out = res;
    if (its_out) *its_out = it;
    if (res > tol) throw Error("ground state: inverse iteration did not converge");
}

// ------------------------------------------------------------------ learned part
// Positional fields of the network at the Greville points of the fine level: 1, V,
// h dV, h^2 d2V (6), V^2, x/L (3), r^2/L^2 for the soft-core potential
// V = -s Z / sqrt(r^2 + a); each channel scaled to unit maximum over the grid.
// fg: the same fields' parameters for their evaluation on the device (without coordinates)
void learned_fields(tdz_ctx *c, std::vector<float> &P, std::vector<double> &scale, tdz::CnnFieldGen &fg)
{
    Level &L = *c->lv[0];
    constexpr int C = tdz::kCnnChannels;
    std::vector<double> gr[3];
    double h[3] = {0, 0, 0}, Lx[3] = {1, 1, 1};
    for (int d = 0; d < 3; ++d) {
        if (d < c->prob.dim) {
            gr[d] = tdz::greville(L.ax[d].t, L.ax[d].p);
            h[d] = (L.ax[d].hi - L.ax[d].lo) / L.ax[d].E;
            Lx[d] = std::max(std::fabs(L.ax[d].lo), std::fabs(L.ax[d].hi));
        } else {
            gr[d] = {0.0};
        }
    }
    const double sZ = c->ph.s * c->ph.pot[0], a = c->ph.pot_kind == 2 ? c->ph.pot[1] : 1.0;
    std::vector<double> f((size_t)L.N * C, 0.0);
    for (int k = 0; k < L.n[2]; ++k)
        for (int j = 0; j < L.n[1]; ++j)
            for (int i = 0; i < L.n[0]; ++i) {
                const double x[3] = {gr[0][i], gr[1][j], gr[2][k]};
                const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
                const double rho = 1.0 / std::sqrt(r2 + a), rho3 = rho * rho * rho, rho5 = rho3 * rho * rho;
                double *o = &f[(size_t)(i + (size_t)L.n[0] * (j + (size_t)L.n[1] * k)) * C];
                const double V = -sZ * rho;
                o[0] = 1.0;
                o[1] = V;
                for (int d = 0; d < 3; ++d) o[2 + d] = sZ * x[d] * rho3 * h[d];
                for (int d = 0; d < 3; ++d) o[5 + d] = sZ * (rho3 - 3.0 * x[d] * x[d] * rho5) * h[d] * h[d];
                o[8] = -3.0 * sZ * x[0] * x[1] * rho5 * h[0] * h[1];
                o[9] = -3.0 * sZ * x[0] * x[2] * rho5 * h[0] * h[2];
                o[10] = -3.0 * sZ * x[1] * x[2] * rho5 * h[1] * h[2];
                o[11] = V * V;
                for (int d = 0; d < 3; ++d) o[12 + d] = x[d] / Lx[d];
                o[15] = r2 / (Lx[0] * Lx[0] + Lx[1] * Lx[1] + Lx[2] * Lx[2]);
            }
    std::vector<double> mx(C, 0.0);
    for (size_t v = 0; v < (size_t)L.N; ++v)
        for (int q = 0; q < C; ++q) mx[q] = std::max(mx[q], std::fabs(f[v * C + q]));
    P.resize(f.size());
    for (size_t v = 0; v < (size_t)L.N; ++v)
        for (int q = 0; q < C; ++q) P[v * C + q] = (float)(mx[q] > 0 ? f[v * C + q] / mx[q] : 0.0);
    scale = mx;
    fg.sZ = (float)sZ;
    fg.a = (float)a;
    for (int d = 0; d < 3; ++d) {
        fg.h[d] = (float)h[d];
        fg.inv_l[d] = (float)(1.0 / Lx[d]);
    }
    fg.inv_r2 = (float)(1.0 / (Lx[0] * Lx[0] + Lx[1] * Lx[1] + Lx[2] * Lx[2]));
    for (int q = 0; q < C; ++q) fg.scale[q] = (float)(mx[q] > 0 ? 1.0 / mx[q] : 0.0);
}

void init_weights(tdz_ctx *c);

void setup_learned(tdz_ctx *c)
{
    LearnState &ls = c->learn;
    if (ls.ready) return;
    setup_mixed(c);
    setup_corrected(c);
    Level &L = *c->lv[0];
    constexpr int C = tdz::kCnnChannels, CI = tdz::kCnnInputs;
    const long N = L.N;
    std::vector<float> P;
    learned_fields(c, P, ls.fscale, ls.fgen);
    ls.P = upload(c, P);
    ls.W1 = track(c, dalloc<float>(27 * CI * C));
    ls.W2 = track(c, dalloc<float>(27 * C * C));
    ls.mb = track(c, dalloc<unsigned short>(tdz::cnn_infer_work_bytes(N) / 2));
    ls.out = track(c, dalloc<float>((size_t)2 * N));
    ls.feat = track(c, dalloc<fcomplex>((size_t)CI * N));
    for (int d = 0; d < 3; ++d) {
        const size_t sz = L.M[d].size();
        ls.Xr[d] = upload(c, to_float(L.Xb[d]));
        std::vector<float> g(L.n[d], 0.f);
        if (d < c->prob.dim) {
            const std::vector<double> gd = tdz::greville(L.ax[d].t, L.ax[d].p);
            for (int i = 0; i < L.n[d]; ++i) g[i] = (float)gd[i];
        }
        ls.gr[d] = upload(c, g);
        ls.fgen.gr[d] = ls.gr[d];
    }
    ls.ready = true;
    init_weights(c);
}

// Initial weights: the first-order Taylor closure of the potential product,
//   V y - V(x_v) M y ~ sum_d dV/dx_d(x_v) (X_d y - x_d M y),
// routed through hidden channel 2 + d (its positional field is h_d dV/dx_d scaled to
// unit maximum) with the current input and target scales, plus small deterministic
// noise on every other weight so that no channel starts without a gradient.
void init_weights(tdz_ctx *c)
{
    LearnState &ls = c->learn;
    Level &L = *c->lv[0];
    constexpr int C = tdz::kCnnChannels, CI = tdz::kCnnInputs;
    ls.hW1.assign(27 * CI * C, 0.f);
    ls.hW2.assign(27 * C * C, 0.f);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return ((st >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
    };
    for (auto &w : ls.hW1) w = (float)(rnd() * 1e-2);
    for (auto &w : ls.hW2) w = (float)(rnd() * 1e-3);
    constexpr int centre = 13;   // tap (0, 0, 0)
    for (int d = 0; d < c->prob.dim; ++d) {
        const double h = (L.ax[d].hi - L.ax[d].lo) / L.ax[d].E;
        ls.hW1[(centre * CI + 1 + d) * C + 2 + d] = 1.f;
        ls.hW2[(centre * C + 2 + d) * C + 0] = (float)(ls.fscale[2 + d] / h * ls.sout / ls.sin.v[1 + d]);
    }
    upload_weights(ls);
}

void setup_training_buffers(tdz_ctx *c)
{
    LearnState &ls = c->learn;
    if (ls.m) return;
    constexpr int C = tdz::kCnnChannels;
    const long N = c->lv[0]->N;
    ls.m = track(c, dalloc<float>((size_t)2 * N * C));
    ls.dov = track(c, dalloc<float>((size_t)2 * N * C));
    ls.dh = track(c, dalloc<float>((size_t)2 * N * C));
    ls.partial = track(c, dalloc<float>((size_t)256 * 27 * C * C));
    ls.dW1 = track(c, dalloc<float>(27 * tdz::kCnnInputs * C));
    ls.lpart = track(c, dalloc<double>(tdz::kLossBlocks));
    ls.dW2 = track(c, dalloc<float>(27 * C * C));
}

// Normalised loss sum_s |N(y_s) - t_s|^2 / |t_s|^2 and its gradient (summed over the
// samples), FP32 forward/backward kernels; gradients returned on the host.
double learned_loss_grad(tdz_ctx *c, std::vector<double> &g1, std::vector<double> &g2, bool want_grad)
{
    LearnState &ls = c->learn;
    Level &L = *c->lv[0];
    constexpr int C = tdz::kCnnChannels;
    const long N = L.N;
    const int n3[3] = {L.n[0], L.n[1], L.n[2]};
    g1.assign(27 * tdz::kCnnInputs * C, 0.0);
    g2.assign(27 * C * C, 0.0);
    std::vector<float> h1(27 * tdz::kCnnInputs * C), h2(27 * C * C);
    double loss = 0;
    const int nchunk = 256;
    for (size_t s = 0; s < ls.ys.size(); ++s) {
        tdz::cnn_forward_f32(c->st, n3, ls.ys[s], ls.W1, ls.W2, ls.P, ls.m, ls.out);
        tdz::cnn_loss(c->st, N, ls.out, ls.ts[s], ls.lpart);
        std::vector<double> lp(tdz::kLossBlocks);
        HIPCHK(hipMemcpyAsync(lp.data(), ls.lpart, lp.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        double l = 0;
        for (double v : lp) l += v;
        loss += l / ls.tnorm2[s];
        if (!want_grad) continue;
        tdz::cnn_backward_f32(c->st, n3, ls.ys[s], ls.ts[s], (float)(1.0 / ls.tnorm2[s]), ls.W2, ls.P, ls.m, ls.out,
                              ls.dov, ls.dh, ls.partial, nchunk, ls.dW1, ls.dW2);
        HIPCHK(hipMemcpyAsync(h1.data(), ls.dW1, h1.size() * 4, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(h2.data(), ls.dW2, h2.size() * 4, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (size_t i = 0; i < h1.size(); ++i) g1[i] += h1[i];
        for (size_t i = 0; i < h2.size(); ++i) g2[i] += h2[i];
    }
    return loss;
}

// Input scales (unit mean square of each input field over the samples) and the target
// scale, applied to the stored samples; inference applies the same scales.
void normalise_samples(tdz_ctx *c)
{
    LearnState &ls = c->learn;
    const long N = c->lv[0]->N;
    constexpr int CI = tdz::kCnnInputs;
    std::vector<fcomplex> h((size_t)CI * N);
    double su[CI] = {0, 0, 0, 0}, stt = 0;
    for (size_t s = 0; s < ls.ys.size(); ++s) {
        HIPCHK(hipMemcpy(h.data(), ls.ys[s], h.size() * sizeof(fcomplex), hipMemcpyDeviceToHost));
        for (int k = 0; k < CI; ++k)
            for (long v = 0; v < N; ++v) {
                const fcomplex z = h[(size_t)k * N + v];
                su[k] += (double)z.re * z.re + (double)z.im * z.im;
            }
        stt += ls.tnorm2[s];
    }
    const double cnt = (double)ls.ys.size() * N;
    for (int k = 0; k < CI; ++k) ls.sin.v[k] = su[k] > 0 ? (float)std::sqrt(cnt / su[k]) : 1.f;
    ls.sout = stt > 0 ? std::sqrt(cnt / stt) : 1.0;
    std::vector<fcomplex> th(N);
    for (size_t s = 0; s < ls.ys.size(); ++s) {
        tdz::scale_inputs(c->st, N, ls.sin, ls.ys[s]);
        HIPCHK(hipMemcpyAsync(th.data(), ls.ts[s], N * sizeof(fcomplex), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (auto &z : th) z = {(float)(z.re * ls.sout), (float)(z.im * ls.sout)};
        HIPCHK(hipMemcpy(ls.ts[s], th.data(), N * sizeof(fcomplex), hipMemcpyHostToDevice));
        ls.tnorm2[s] *= ls.sout * ls.sout;
    }
    HIPCHK(hipStreamSynchronize(c->st));
}

// Adam on the harvested samples.
void learned_fit(tdz_ctx *c, int iters, double lr, double *loss0, double *loss1)
{
    LearnState &ls = c->learn;
    setup_training_buffers(c);
    std::vector<double> g1, g2, m1(ls.hW1.size(), 0.0), v1(ls.hW1.size(), 0.0), m2(ls.hW2.size(), 0.0),
        v2(ls.hW2.size(), 0.0);
    const double b1 = 0.9, b2 = 0.999, eps = 1e-12;
    double first = -1, last = 0;
    for (int it = 1; it <= iters; ++it) {
        last = learned_loss_grad(c, g1, g2, true);
        if (first < 0) first = last;
        const double c1 = 1.0 - std::pow(b1, it), c2 = 1.0 - std::pow(b2, it);
        auto step = [&](std::vector<float> &w, const std::vector<double> &g, std::vector<double> &m,
                        std::vector<double> &v) {
            for (size_t i = 0; i < w.size(); ++i) {
                m[i] = b1 * m[i] + (1 - b1) * g[i];
                v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
                w[i] -= (float)(lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps));
            }
        };
        step(ls.hW1, g1, m1, v1);
        step(ls.hW2, g2, m2, v2);
        upload_weights(ls);
    }
    last = learned_loss_grad(c, g1, g2, false);
    // reported per sample: 1 is the loss of a network that outputs zero
    const double ns = (double)ls.ys.size();
    if (loss0) *loss0 = first / ns;
    if (loss1) *loss1 = last / ns;
}

}  // namespace

extern "C" {

int tdz_project_gaussian(tdz_ctx *c, const double centre[3], const double width[3], const double momentum[3],
                         double *psi)
{
    return guarded([&] { project_gaussian(c, centre, width, momentum, reinterpret_cast<dcomplex *>(psi)); });
}

int tdz_ground_state(tdz_ctx *c, double *psi, double sigma, double tol, int maxit, double *energy,
                     double *residual, int *iterations)
{
    return guarded([&] {
        ground_state(c, reinterpret_cast<dcomplex *>(psi), sigma, tol, maxit, energy, residual, iterations);
        HIPCHK(hipStreamSynchronize(c->st));   // the results are final when the call returns
    });
}

int tdz_step_pade22(tdz_ctx *c, double *psi, double h, const double sE[3], int slot, double rtol, int maxit,
                    tdz_ksp_stats *stats)
{
    return guarded([&] {
        if (slot < 0 || slot > 1) throw Error("slot must be 0 or 1");
        pade22_step(c, reinterpret_cast<dcomplex *>(psi), h, sE, slot, rtol, maxit, stats);
        HIPCHK(hipStreamSynchronize(c->st));   // the results are final when the call returns
    });
}

int tdz_evolve_rational(tdz_ctx *c, double *psi, double t0, double dt, int nsteps, const tdz_pulse *pulse,
                        int magnus, double rtol, int maxit, double *obs, int *its)
{
    return guarded([&] {
        dcomplex *x = reinterpret_cast<dcomplex *>(psi);
        // commutator-free Magnus 4 (src/propagator.cpp:729-758): Gauss nodes c1,2, weights a1,2
        const double s3 = std::sqrt(3.0);
        if (!c->pc_ready) pc_setup(c, c->pc_kind);
        multigrid_checks(c, pulse, magnus ? dt / (4.0 * s3) : dt / (2.0 * s3), dt);   // dt/(2 sqrt12), dt/sqrt12
        const double c1 = 0.5 - s3 / 6.0, c2 = 0.5 + s3 / 6.0;
        const double a1 = (3.0 - 2.0 * s3) / 12.0, a2 = (3.0 + 2.0 * s3) / 12.0;
        for (int s = 0; s < nsteps; ++s) {
            const double t = t0 + s * dt;
            tdz_ksp_stats st{};
            int total = 0;
            if (!magnus) {
                double E[3] = {0, 0, 0};
                if (pulse) tdz_pulse_coupling(c, pulse, t + 0.5 * dt, E);
                pade22_step(c, x, dt, E, 0, rtol, maxit, &st);
                total = st.its;
            } else {
                double E1[3] = {0, 0, 0}, E2[3] = {0, 0, 0}, Ea[3], Eb[3];
                if (pulse) {
                    tdz_pulse_coupling(c, pulse, t + c1 * dt, E1);
                    tdz_pulse_coupling(c, pulse, t + c2 * dt, E2);
                }
                for (int a = 0; a < 3; ++a) {
                    // the first half step (a2 > |a1|) weights the earlier Gauss node most, the second the later one
                    Ea[a] = 2.0 * (a2 * E1[a] + a1 * E2[a]);
                    Eb[a] = 2.0 * (a1 * E1[a] + a2 * E2[a]);
                }
                pade22_step(c, x, 0.5 * dt, Ea, 0, rtol, maxit, &st);
                total = st.its;
                pade22_step(c, x, 0.5 * dt, Eb, 1, rtol, maxit, &st);
                total += st.its;
            }
            if (its) its[s] = total;
            if (obs) {
                double o[6];
                if (tdz_observables(c, psi, o)) throw Error(g_err);
                obs[7 * s] = t + dt;
                for (int k = 0; k < 6; ++k) obs[7 * s + 1 + k] = o[k];
            }
        }
        HIPCHK(hipStreamSynchronize(c->st));   // the results are final when the call returns
    });
}

int tdz_pulse_coupling(const tdz_ctx *c, const tdz_pulse *pulse, double t, double sE[3])
{
    if (!c || !pulse || !sE) { g_err = "null argument"; return 1; }
    for (int a = 0; a < 3; ++a) sE[a] = a < c->prob.dim ? c->ph.s * pulse_field(*pulse, a, t) : 0.0;
    return 0;
}

}  // extern "C"

namespace {
// psi^H T psi for T = M, H (K + sV + CAP, no field), D_x, D_y, D_z, K, sV, in that order
// (v[0..6]), and v[7] = psi0^H M psi when psi0 is given (the autocorrelation)
// (MonitorProducts, src/monitor.cpp:1145-1194). Vectors are reduced on the host in FP64.
void expectations(tdz_ctx *c, const double *psi, const double *psi0, cplx *v)
{
    Level &L = *c->lv[0];
    const long N = L.N;
    std::vector<dcomplex> hp(N), hv(N), h0;
    HIPCHK(hipMemcpyAsync(hp.data(), psi, N * sizeof(dcomplex), hipMemcpyDeviceToHost, c->st));
    if (psi0) {
        h0.resize(N);
        HIPCHK(hipMemcpyAsync(h0.data(), psi0, N * sizeof(dcomplex), hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    auto dot = [&](int term_or_h, const std::vector<dcomplex> &left) {
        if (term_or_h >= 0) {
            if (tdz_term_apply(c, term_or_h, psi, reinterpret_cast<double *>(c->sd.w1))) throw Error(g_err);
        } else {
            const cplx a0 = c->alpha, g0 = c->g;
            double s0[3] = {c->sE[0], c->sE[1], c->sE[2]};
            const double zero[3] = {0, 0, 0};
            set_shift(c, 0.0, 1.0, zero);
            level_apply(c, L, reinterpret_cast<const dcomplex *>(psi), c->sd.w1);
            set_shift(c, a0, g0, s0);
        }
        HIPCHK(hipMemcpyAsync(hv.data(), c->sd.w1, N * sizeof(dcomplex), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        cplx s = 0;
        for (long i = 0; i < N; ++i) s += std::conj(cplx(left[i].re, left[i].im)) * cplx(hv[i].re, hv[i].im);
        return s;
    };
    v[0] = dot(TDZ_TERM_M, hp);
    v[1] = dot(-1, hp);
    for (int a = 0; a < 3; ++a) v[2 + a] = a < c->prob.dim ? dot(TDZ_TERM_DX + a, hp) : cplx(0);
    v[5] = dot(TDZ_TERM_K, hp);
    v[6] = c->ph.pot_kind ? dot(TDZ_TERM_V, hp) : cplx(0);
    v[7] = psi0 ? dot(TDZ_TERM_M, h0) : cplx(0);
}
}  // namespace

extern "C" {

}  // extern "C"

namespace tdz {
int fail(const std::string &msg)
{
    g_err = msg;
    return 1;
}

int ctx_device(const tdz_ctx *c) { return c ? c->device : -1; }

// The observables of tdz_observables on the device, with the reductions' partial sums and
// results in the caller's scratch scr (noalloc_scratch_len values, device memory): no
// allocation, no host staging. The products are those of tdz_term_apply (wall rows: the
// identity) and of the shifted operator with alpha = 0, g = 1 (H without field).
int observables_device(tdz_ctx *c, const double *psi, double out[6], dcomplex *scr)
{
    return guarded([&] {
        if (!c || !psi || !out || !scr) throw Error("null argument");
        if (c->layered) throw Error("tdz_observables_noalloc: layered structures are not supported");
        Level &L = *c->lv[0];
        const long N = L.N;
        if (c->ph.pot_kind && L.simple && !c->pot_scr) throw Error("potential scratch missing");
        const dcomplex *x = reinterpret_cast<const dcomplex *>(psi);
        dcomplex *w = c->sd.w1, *res = scr + (long)kObsDots * kRedBlocks;
        const unsigned char *wall[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
        auto dot = [&](int slot) {   // res[slot] = <psi, w>; the partials reuse the front of scr
            tdz::launch_multi_dot(c->st, x, N, 1, w, N, scr, res + slot);
        };
        mass_apply(c, x, w);
        dot(0);
        const cplx a0 = c->alpha, g0 = c->g;
        const double s0[3] = {c->sE[0], c->sE[1], c->sE[2]}, zero[3] = {0, 0, 0};
        set_shift(c, 0.0, 1.0, zero);
        level_apply(c, L, x, w);
        set_shift(c, a0, g0, s0);
        dot(1);
        for (int a = 0; a < c->prob.dim; ++a) {
            tdz::SepBandsT<double> db = sep_bands<double>(L);
            for (int d = 0; d < 3; ++d) db.B[d] = d == a ? c->pred.Xband[a] : c->pred.Zband[d];
            tdz::launch_sep_apply<double>(c->st, db, x, w, c->sd.tmp);
            if (c->prob.dirichlet) tdz::launch_wall_fix<double>(c->st, L.n, wall, x, w);
            dot(2 + a);
        }
        dcomplex h[kObsDots] = {};
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h, res, (2 + c->prob.dim) * sizeof(dcomplex), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        out[0] = h[0].re;
        out[1] = h[1].re;
        out[2] = h[1].im;
        for (int a = 0; a < 3; ++a) out[3 + a] = a < c->prob.dim ? h[2 + a].re : 0.0;
    });
}
}  // namespace tdz

extern "C" {

int tdz_observables(tdz_ctx *c, const double *psi, double out[6])
{
    return guarded([&] {
        cplx v[8];   // expectations() fills 8 values
        expectations(c, psi, nullptr, v);
        out[0] = v[0].real();
        out[1] = v[1].real();
        out[2] = v[1].imag();
        for (int a = 0; a < 3; ++a) out[3 + a] = v[2 + a].real();
    });
}

int tdz_pade_history(tdz_ctx *c, int set, int counts[4], double *hist)
{
    return guarded([&] {
        const long N = c->lv[0]->N;
        if (!set) {
            for (int s = 0; s < 4; ++s) counts[s] = c->hist ? c->hist_count[s] : 0;
            if (c->hist && hist)
                HIPCHK(hipMemcpy(hist, c->hist, (size_t)4 * c->hist_order * N * sizeof(dcomplex),
                                 hipMemcpyDeviceToHost));
            else if (hist)
                std::memset(hist, 0, (size_t)4 * c->guess_order * N * sizeof(dcomplex));
            return;
        }
        pade_hist_alloc(c);
        for (int s = 0; s < 4; ++s) c->hist_count[s] = std::max(0, std::min(counts[s], c->guess_order));
        if (hist)
            HIPCHK(hipMemcpy(c->hist, hist, (size_t)4 * c->guess_order * N * sizeof(dcomplex), hipMemcpyHostToDevice));
    });
}

int tdz_set_pade_guess(tdz_ctx *c, int extrapolate, int order)
{
    return guarded([&] {
        if (order < 1 || order > 5) throw Error("Pade guess order must be 1..5");
        c->guess_extrap = extrapolate != 0;
        if (order != c->guess_order) {
            c->guess_order = order;
            for (int s = 0; s < 4; ++s) c->hist_count[s] = 0;
        }
    });
}

int tdz_trace_row(tdz_ctx *c, const double *psi, const double *psi0, double t, const double E[3], double *dip,
                  double *energy, double *ac)
{
    return guarded([&] {
        cplx v[8];
        expectations(c, psi, psi0, v);
        const int dim = c->prob.dim;
        const double q = -c->ph.s;   // charge
        if (dip) {
            dip[0] = t;
            for (int a = 0; a < dim; ++a) {
                dip[1 + a] = E ? E[a] : 0.0;
                dip[1 + dim + a] = q * v[2 + a].real();
                dip[1 + 2 * dim + a] = 0.0;   // acceleration is not computed
            }
        }
        if (energy) {
            double inter = 0;
            for (int a = 0; a < dim; ++a) inter += c->ph.s * (E ? E[a] : 0.0) * v[2 + a].real();
            energy[0] = t;
            energy[1] = v[5].real();                // kinetic
            energy[2] = v[6].real();                // potential s V
            energy[3] = inter;                      // s E . <D>
            energy[4] = v[1].real() + inter;        // <H> + interaction
            energy[5] = v[0].real() / c->ph.mass;   // <1/m> (constant mass)
            energy[6] = v[0].real();                // norm
        }
        if (ac) {
            ac[0] = t;
            ac[1] = v[7].real();
            ac[2] = v[7].imag();
        }
    });
}

int tdz_evolve_cn(tdz_ctx *c, double *psi, double t0, double dt, int nsteps, const tdz_pulse *pulse, double rtol,
                  int maxit, double *obs, int *its)
{
    return guarded([&] {
        if (!c->pc_ready) pc_setup(c, c->pc_kind);
        multigrid_checks(c, pulse, 0.5 * dt, dt);
        dcomplex *x = reinterpret_cast<dcomplex *>(psi);
        if (!pred_continues(c, x, dt)) c->pred.count = 0;
        for (int s = 0; s < nsteps; ++s) {
            const double tm = t0 + (s + 0.5) * dt;
            double sE[3] = {0, 0, 0};
            if (pulse) tdz_pulse_coupling(c, pulse, tm, sE);
            tdz_ksp_stats st{};
            // order 0 is the current state as the guess (TSTHETA's): the same fused
            // right-hand side and history products, and no unchained first step per call
            cn_step_pred(c, x, dt, sE, rtol, maxit, &st);
            // the separable solves depend on (alpha, g) only and every level operator follows
            // the field through set_shift, so a rebuild has nothing to redo but is counted
            if (refresh_due(c, c->refresh_base, st.its)) pc_setup(c, c->pc_kind);
            if (its) its[s] = st.its;
            if (obs) {
                double o[6];
                if (tdz_observables(c, psi, o)) throw Error(g_err);
                obs[7 * s] = t0 + (s + 1) * dt;
                for (int k = 0; k < 6; ++k) obs[7 * s + 1 + k] = o[k];
            }
        }
        HIPCHK(hipStreamSynchronize(c->st));   // the results are final when the call returns
    });
}

int tdz_learn_train(tdz_ctx *c, const double *psi, double t0, double dt, const tdz_pulse *pulse, int harvest_steps,
                    int max_samples, int iters, double lr, double rtol, double *loss_before, double *loss_after)
{
    return guarded([&] {
        if (c->ph.pot_kind == 0) throw Error("no potential: nothing to learn");
        setup_learned(c);
        LearnState &ls = c->learn;
        Level &L = *c->lv[0];
        const long N = L.N;
        const int kind0 = c->pc_kind;
        // a new training run starts from fresh samples and unit scales
        for (fcomplex *p : ls.ys) untrack(c, p);
        for (fcomplex *p : ls.ts) untrack(c, p);
        ls.ys.clear();
        ls.ts.clear();
        ls.tnorm2.clear();
        ls.sin = tdz::CnnScales{{1.f, 1.f, 1.f, 1.f}};
        ls.sout = 1.0;
        // harvest from a copy of psi with the preconditioner the network augments
        dcomplex *x = dalloc<dcomplex>(N);
        HIPCHK(hipMemcpy(x, psi, N * sizeof(dcomplex), hipMemcpyDeviceToDevice));
        pc_setup(c, TDZ_PC_CORRECTED);
        ls.max_samples = max_samples;
        ls.harvesting = true;
        int base = 0;
        try {
            for (int s = 0; s < harvest_steps; ++s) {
                double sE[3] = {0, 0, 0};
                if (pulse) tdz_pulse_coupling(c, pulse, t0 + (s + 0.5) * dt, sE);
                tdz_ksp_stats st{};
                if (tdz_step_cn(c, reinterpret_cast<double *>(x), dt, sE, rtol, 1000, &st)) throw Error(g_err);
                base = std::max(base, st.its);
            }
        } catch (...) {
            ls.harvesting = false;
            hipFree(x);
            throw;
        }
        ls.harvesting = false;
        HIPCHK(hipFree(x));
        random_samples(c, ls.random_count, ls.seed);
        ls.base_its = base;
        if (ls.ys.empty()) throw Error("no samples harvested");
        normalise_samples(c);
        init_weights(c);
        learned_fit(c, iters, lr, loss_before, loss_after);
        ls.trained = true;
        ls.fallen = false;
        pc_setup(c, kind0 == TDZ_PC_NONE ? TDZ_PC_LEARNED : kind0);
    });
}

int tdz_learn_config(tdz_ctx *c, int random_samples, unsigned long long seed)
{
    return guarded([&] {
        if (random_samples < 0 || random_samples > 64) throw Error("random samples must be 0..64");
        c->learn.random_count = random_samples;
        c->learn.seed = seed;
    });
}

int tdz_learn_weights(tdz_ctx *c, float *W1, float *W2, int set)
{
    return guarded([&] {
        setup_learned(c);
        LearnState &ls = c->learn;
        if (set) {
            std::copy(W1, W1 + ls.hW1.size(), ls.hW1.begin());
            std::copy(W2, W2 + ls.hW2.size(), ls.hW2.begin());
            upload_weights(ls);
            ls.trained = true;
        } else {
            std::copy(ls.hW1.begin(), ls.hW1.end(), W1);
            std::copy(ls.hW2.begin(), ls.hW2.end(), W2);
        }
    });
}

int tdz_learn_fields(tdz_ctx *c, float *P)
{
    return guarded([&] {
        setup_learned(c);
        HIPCHK(hipMemcpy(P, c->learn.P, (size_t)c->lv[0]->N * tdz::kCnnChannels * 4, hipMemcpyDeviceToHost));
    });
}

int tdz_learn_features(tdz_ctx *c, const double *y, double *feat)
{
    return guarded([&] {
        setup_learned(c);
        Level &L = *c->lv[0];
        const long N = L.N;
        constexpr int CI = tdz::kCnnInputs;
        tdz::launch_demote(c->st, reinterpret_cast<const dcomplex *>(y), L.f.z, N);
        fcomplex *tmp[7] = {c->sf.tmp[0], c->sf.tmp[1], c->sf.tmp[2], c->sf.tmp[3], c->sf.w1, c->sf.w2, c->sf.u2};
        const tdz::CnnScales unit{{1.f, 1.f, 1.f, 1.f}};
        tdz::fused_products(c->st, fused_bands(c), unit, L.f.z, tmp, L.f.t, c->sf.u1, c->learn.feat);
        std::vector<fcomplex> h((size_t)CI * N);
        HIPCHK(hipMemcpyAsync(h.data(), c->learn.feat, h.size() * sizeof(fcomplex), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (size_t k = 0; k < h.size(); ++k) {
            feat[2 * k] = h[k].re;
            feat[2 * k + 1] = h[k].im;
        }
    });
}

int tdz_learn_eval(tdz_ctx *c, const double *u, const double *target, double *loss, float *dW1, float *dW2,
                   double *out_f32, double *out_mfma)
{
    return guarded([&] {
        setup_learned(c);
        setup_training_buffers(c);
        LearnState &ls = c->learn;
        Level &L = *c->lv[0];
        const long N = L.N;
        constexpr int CI = tdz::kCnnInputs;
        std::vector<fcomplex> uh((size_t)CI * N), th(N);
        for (size_t k = 0; k < uh.size(); ++k) uh[k] = {(float)u[2 * k], (float)u[2 * k + 1]};
        double t2 = 0;
        for (long k = 0; k < N; ++k) {
            th[k] = {(float)target[2 * k], (float)target[2 * k + 1]};
            t2 += (double)th[k].re * th[k].re + (double)th[k].im * th[k].im;
        }
        fcomplex *y32 = upload(c, uh), *t32 = upload(c, th);
        auto ys = ls.ys, ts = ls.ts;
        auto tn = ls.tnorm2;
        ls.ys = {y32};
        ls.ts = {t32};
        ls.tnorm2 = {t2 > 0 ? t2 : 1.0};
        std::vector<double> g1, g2;
        const double l = learned_loss_grad(c, g1, g2, true);
        ls.ys = ys;
        ls.ts = ts;
        ls.tnorm2 = tn;
        if (loss) *loss = l;
        if (dW1) for (size_t i = 0; i < g1.size(); ++i) dW1[i] = (float)g1[i];
        if (dW2) for (size_t i = 0; i < g2.size(); ++i) dW2[i] = (float)g2[i];
        std::vector<float> o((size_t)2 * N);
        if (out_f32) {
            HIPCHK(hipMemcpy(o.data(), ls.out, o.size() * 4, hipMemcpyDeviceToHost));   // forward of the loss
            for (size_t i = 0; i < o.size(); ++i) out_f32[i] = o[i];
        }
        if (out_mfma) {
            const int n3[3] = {L.n[0], L.n[1], L.n[2]};
            tdz::cnn_infer_mfma(c->st, n3, y32, ls.W1, ls.W2, ls.fgen, ls.mb, ls.out, fcomplex{1.f, 0.f}, nullptr, true);
            ls.packed_ver = ls.wver;
            HIPCHK(hipStreamSynchronize(c->st));
            HIPCHK(hipMemcpy(o.data(), ls.out, o.size() * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < o.size(); ++i) out_mfma[i] = o[i];
        }
        untrack(c, y32);
        untrack(c, t32);
    });
}

int tdz_learn_status(tdz_ctx *c, int *base_its, int *fallbacks, int *active)
{
    if (!c) { g_err = "null context"; return 1; }
    if (base_its) *base_its = c->learn.base_its;
    if (fallbacks) *fallbacks = c->learn.fallbacks;
    if (active) *active = c->pc_kind == TDZ_PC_LEARNED && c->learn.trained && !c->learn.fallen;
    return 0;
}

int tdz_pc_sweep_stats(tdz_ctx *c, long *counts)
{
    if (!c || !counts) { g_err = "null argument"; return 1; }
    for (int net = 0; net < 2; ++net)
        for (int k = 0; k < 4; ++k) counts[4 * net + k] = c->sweep_count[net][k];
    return 0;
}

int tdz_timing(tdz_ctx *c, int enable)
{
    return guarded([&] {
        HIPCHK(hipStreamSynchronize(c->st));
        c->timing = enable != 0;
        c->ev_used = 0;
    });
}

int tdz_timing_stats(tdz_ctx *c, double *total_ms, long *launches)
{
    return guarded([&] {
        HIPCHK(hipStreamSynchronize(c->st));
        double t = 0;
        for (size_t i = 0; i < c->ev_used; ++i) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[i].first, c->ev[i].second));
            t += ms;
        }
        *total_ms = t;
        *launches = (long)c->ev_used;
    });
}

}  // extern "C"

// ===================================================================== batched scan
namespace {
// Device data of a run for captured CN steps: the static part of each axis band
// T0 = S/(2m) - i W (complex), the position bands, the field s E(t_n + dt/2) of every
// step, and the work of the device-side FGMRES(k).
void setup_capture(tdz_ctx *c, double dt, const std::vector<double> &sE, int nsteps, int k)
{
    Level &L = *c->lv[0];
    auto &C = c->cap;
    if (k < 1 || k > tdz::kGmMax) throw Error("captured FGMRES length must be in [1, 8]");
    if (k > c->restart) throw Error("captured FGMRES length exceeds the Krylov basis");
    C.k = k;
    C.dt = dt;
    C.nsteps = nsteps;
    C.gm = track(c, dalloc<tdz::GmState>(1));
    C.bnorm2 = track(c, dalloc<double>(1));
    C.coef = track(c, dalloc<dcomplex>(tdz::kGmMax));
    C.step = track(c, dalloc<int>(1));
    C.res = track(c, dalloc<double>(nsteps + 1));
    HIPCHK(hipMemset(C.step, 0, sizeof(int)));
    HIPCHK(hipMemset(C.res, 0, (nsteps + 1) * sizeof(double)));
    C.sE = upload(c, sE);
    const double inv2m = 0.5 / c->ph.mass;
    for (int d = 0; d < 3; ++d) {
        const size_t sz = L.M[d].size();
        std::vector<dcomplex> t0(sz, dcomplex{0, 0});
        std::vector<double> xb(sz, 0.0);
        if (d < c->prob.dim)
            for (size_t q = 0; q < sz; ++q) {
                t0[q] = {inv2m * L.S[d][q], c->cap_on ? -L.Wc[d][q] : 0.0};
                xb[q] = L.Xb[d][q];
            }
        C.T0[d] = upload(c, t0);
        C.Xb[d] = upload(c, xb);
        C.bandR[d] = track(c, dalloc<dcomplex>(sz));
        C.bandL[d] = track(c, dalloc<dcomplex>(sz));
    }
    C.M0 = upload(c, L.M[0]);
}

// Records one CN step of psi into the context's stream (which is capturing): field
// bands of this step, rhs = (M - g H_E) psi, r0 = 2 (rhs - M psi), k FGMRES iterations
// with the Hessenberg problem on the device, the solution update and the step counter.
// No host synchronisation and no memset inside.
void record_cn_step(tdz_ctx *c, dcomplex *x)
{
    Level &L = *c->lv[0];
    auto &C = c->cap;
    const long N = L.N;
    const int k = C.k;
    const cplx g(0.0, 0.5 * C.dt);
    for (int d = 0; d < 3; ++d) {
        const long sz = (long)L.M[d].size();
        tdz::launch_field_band(c->st, sz, d == 0 ? C.M0 : nullptr, C.T0[d], C.Xb[d], dc(-g), C.sE, C.step, d,
                               C.bandR[d]);
        tdz::launch_field_band(c->st, sz, d == 0 ? C.M0 : nullptr, C.T0[d], C.Xb[d], dc(g), C.sE, C.step, d,
                               C.bandL[d]);
    }
    dcomplex *own[3] = {L.d.B[0], L.d.B[1], L.d.B[2]};
    const cplx g0 = c->g;
    for (int d = 0; d < 3; ++d) L.d.B[d] = C.bandR[d];
    c->g = -g;
    level_apply(c, L, x, c->rhs);
    dcomplex *V0 = c->Vb;
    if (C.pred > 0) {
        // guess: the extrapolation from this state and the two before (as many as the run
        // has taken steps), then r0 = rhs - (M + g H) guess by one operator apply
        tdz::launch_extrapolate(c->st, C.step, C.pred, x, C.h1, C.h2, c->sd.w2, N);
        HIPCHK(hipMemcpyAsync(C.h2, C.h1, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(C.h1, x, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(x, c->sd.w2, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
        for (int d = 0; d < 3; ++d) L.d.B[d] = C.bandL[d];
        c->g = g;
        level_apply(c, L, x, V0);
        tdz::launch_axpby(c->st, {1.0, 0}, c->rhs, {-1.0, 0}, V0, N);
    } else {   // r0 = 2 (rhs - M x): the guess is the state
        tdz::SepBandsT<double> mb = sep_bands<double>(L);
        for (int d = 0; d < 3; ++d) mb.B[d] = L.d.Bmass[d];
        tdz::launch_sep_apply<double>(c->st, mb, x, c->sd.w1, c->sd.tmp);
        const unsigned char *wall[3] = {L.dWall[0], L.dWall[1], L.dWall[2]};
        if (c->prob.dirichlet) tdz::launch_wall_fix<double>(c->st, L.n, wall, x, c->sd.w1);
        HIPCHK(hipMemcpyAsync(V0, c->rhs, N * sizeof(dcomplex), hipMemcpyDeviceToDevice, c->st));
        tdz::launch_axpby(c->st, {-2.0, 0}, c->sd.w1, {2.0, 0}, V0, N);
    }
    tdz::launch_norm2(c->st, c->rhs, N, c->dpartr, C.bnorm2);
    tdz::launch_norm2(c->st, V0, N, c->dpartr, c->dnorm);
    tdz::launch_gm_start(c->st, c->dnorm, C.bnorm2, C.gm);
    tdz::launch_scale_invsqrt(c->st, V0, c->dnorm, V0, N);
    for (int d = 0; d < 3; ++d) L.d.B[d] = C.bandL[d];
    c->g = g;
    for (int j = 0; j < k; ++j) {
        dcomplex *Vj = c->Vb + (long)j * N, *Zj = c->Zb + (long)j * N, *w = c->Vb + (long)(j + 1) * N;
        pc_apply_dev(c, Vj, Zj);
        level_apply(c, L, Zj, w);
        tdz::launch_multi_dot(c->st, c->Vb, N, j + 1, w, N, c->dpart, c->dh);
        tdz::launch_multi_axpy_norm(c->st, c->Vb, N, j + 1, c->dh, w, N, c->dpartr, c->dnorm);
        tdz::launch_multi_dot(c->st, c->Vb, N, j + 1, w, N, c->dpart, c->dh2);
        tdz::launch_multi_axpy_norm(c->st, c->Vb, N, j + 1, c->dh2, w, N, c->dpartr, c->dnorm);
        tdz::launch_gm_arnoldi(c->st, j, k, c->dh, c->dh2, c->dnorm, C.gm);
        tdz::launch_scale_invsqrt(c->st, w, c->dnorm, w, N);
    }
    tdz::launch_gm_solve(c->st, k, k, C.gm, C.coef, C.res, C.step);
    tdz::launch_multi_axpy(c->st, c->Zb, N, k, C.coef, x, N);
    tdz::launch_step_advance(c->st, C.step);
    for (int d = 0; d < 3; ++d) L.d.B[d] = own[d];
    c->g = g0;
}
}  // namespace

struct tdz_batch {
    std::vector<tdz_ctx *> runs;
    std::vector<dcomplex *> psi;
    std::vector<int> step;    // host copy of each run's step counter
    std::vector<std::vector<double>> sE;
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double dt = 0;
    int nsteps = 0, k = 0;
    long fixups = 0;
    double *hres = nullptr;   // pinned: every run's residual estimate of the last step
};

namespace {
void batch_free(tdz_batch *b)
{
    if (b->st) hipStreamSynchronize(b->st);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (auto e : b->join) hipEventDestroy(e);
    if (b->fork) hipEventDestroy(b->fork);
    for (tdz_ctx *c : b->runs) tdz_destroy(c);
    if (b->hres) hipHostFree(b->hres);
    if (b->st) hipStreamDestroy(b->st);
    delete b;
}

// One graph: fork from the batch stream to every run's stream, one recorded CN step
// per run, join back. The runs are independent branches of the graph.
void capture_batch(tdz_batch *b)
{
    HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    try {
        HIPCHK(hipEventRecord(b->fork, b->st));
        for (size_t r = 0; r < b->runs.size(); ++r) {
            tdz_ctx *c = b->runs[r];
            HIPCHK(hipStreamWaitEvent(c->st, b->fork, 0));
            record_cn_step(c, b->psi[r]);
            HIPCHK(hipEventRecord(b->join[r], c->st));
            HIPCHK(hipStreamWaitEvent(b->st, b->join[r], 0));
        }
    } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(b->st, &g);
        if (g) hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(b->st, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}
}  // namespace

extern "C" {

int tdz_batch_create(const tdz_problem *probs, const tdz_pulse *pulses, int nruns, double dt, int nsteps,
                     int pc_kind, int kiters, tdz_batch **out)
{
    if (!probs || !out || nruns < 1 || nsteps < 1 || dt <= 0) { g_err = "bad batch arguments"; return 1; }
    auto *b = new tdz_batch();
    const int rc = guarded([&] {
        if (pc_kind != TDZ_PC_SEPARABLE && pc_kind != TDZ_PC_SEPARABLE_MIXED && pc_kind != TDZ_PC_NONE)
            throw Error("the captured step takes the separable preconditioners (fp64 or complex64) or none");
        b->dt = dt;
        b->nsteps = nsteps;
        b->k = kiters;
        HIPCHK(hipStreamCreate(&b->st));
        HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        for (int r = 0; r < nruns; ++r) {
            tdz_ctx *c = nullptr;
            if (tdz_create(&probs[r], &c)) throw Error(std::string("run ") + std::to_string(r) + ": " + g_err);
            b->runs.push_back(c);
            pc_setup(c, pc_kind);
            dcomplex *p = track(c, dalloc<dcomplex>(c->lv[0]->N));
            HIPCHK(hipMemset(p, 0, c->lv[0]->N * sizeof(dcomplex)));
            b->psi.push_back(p);
            std::vector<double> e((size_t)3 * (nsteps + 1), 0.0);
            for (int s = 0; s <= nsteps; ++s)
                if (pulses) tdz_pulse_coupling(c, &pulses[r], (s + 0.5) * dt, &e[3 * (size_t)s]);
            for (int s = 0; s <= nsteps; ++s)
                for (int a = c->prob.dim; a < 3; ++a) e[3 * (size_t)s + a] = 0.0;
            b->sE.push_back(e);
            setup_capture(c, dt, e, nsteps, kiters);
            hipEvent_t j;
            HIPCHK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
            b->join.push_back(j);
        }
        b->step.assign(nruns, 0);
        HIPCHK(hipHostMalloc((void **)&b->hres, nruns * sizeof(double), hipHostMallocDefault));
        for (tdz_ctx *c : b->runs) HIPCHK(hipStreamSynchronize(c->st));
    });
    if (rc) {
        batch_free(b);
        *out = nullptr;
        return rc;
    }
    *out = b;
    return 0;
}

int tdz_batch_set_predictor(tdz_batch *b, int order)
{
    return guarded([&] {
        if (!b) throw Error("null batch");
        if (order < 0 || order > 2) throw Error("CN predictor order must be 0, 1 or 2");
        if (b->exec) throw Error("batch: the step is already captured (set the predictor before the first evolve)");
        for (tdz_ctx *c : b->runs) {
            auto &C = c->cap;
            C.pred = order;
            if (order > 0 && !C.h1) {
                const long N = c->lv[0]->N;
                C.h1 = track(c, dalloc<dcomplex>(N));
                C.h2 = track(c, dalloc<dcomplex>(N));
                HIPCHK(hipMemsetAsync(C.h1, 0, N * sizeof(dcomplex), c->st));
                HIPCHK(hipMemsetAsync(C.h2, 0, N * sizeof(dcomplex), c->st));
                HIPCHK(hipStreamSynchronize(c->st));
            }
        }
    });
}

int tdz_batch_destroy(tdz_batch *b)
{
    if (b) batch_free(b);
    return 0;
}

int tdz_batch_context(tdz_batch *b, int run, tdz_ctx **ctx, double **psi)
{
    return guarded([&] {
        if (!b || run < 0 || run >= (int)b->runs.size()) throw Error("no such run");
        if (ctx) *ctx = b->runs[run];
        if (psi) *psi = reinterpret_cast<double *>(b->psi[run]);
    });
}

int tdz_batch_evolve(tdz_batch *b, int nsteps, double rtol, int maxit, int *its, double *res)
{
    return guarded([&] {
        const int R = (int)b->runs.size();
        if (b->step[0] + nsteps > b->nsteps) throw Error("batch: past the last precomputed field step");
        for (tdz_ctx *c : b->runs) HIPCHK(hipStreamSynchronize(c->st));
        if (!b->exec) capture_batch(b);
        double *r = b->hres;
        for (int s = 0; s < nsteps; ++s) {
            HIPCHK(hipGraphLaunch(b->exec, b->st));
            for (int q = 0; q < R; ++q)
                HIPCHK(hipMemcpyAsync(&r[q], b->runs[q]->cap.res + b->step[q], sizeof(double),
                                      hipMemcpyDeviceToHost, b->st));
            HIPCHK(hipStreamSynchronize(b->st));
            for (int q = 0; q < R; ++q) {
                tdz_ctx *c = b->runs[q];
                int it = b->k;
                if (!(r[q] <= rtol)) {
                    // the captured cycle fell short: finish this run's step on the host solver
                    // from the graph's iterate and right-hand side, with this step's field
                    const double *e = &b->sE[q][3 * (size_t)b->step[q]];
                    set_shift(c, 1.0, cplx(0.0, 0.5 * b->dt), e);
                    tdz_ksp_stats st{};
                    fgmres(c, c->rhs, b->psi[q], rtol, maxit, &st);
                    // fgmres leaves its final update queued on the run's stream; the next
                    // graph launch (batch stream) reads psi and rewrites the solver's
                    // buffers, so the run's stream must drain first
                    HIPCHK(hipStreamSynchronize(c->st));
                    it += st.its;
                    r[q] = st.res;
                    ++b->fixups;
                }
                if (its) its[(size_t)s * R + q] = it;
                if (res) res[(size_t)s * R + q] = r[q];
                ++b->step[q];
            }
        }
    });
}

int tdz_batch_fixups(const tdz_batch *b, long *fixups)
{
    if (!b || !fixups) { g_err = "null argument"; return 1; }
    *fixups = b->fixups;
    return 0;
}

}  // extern "C"
