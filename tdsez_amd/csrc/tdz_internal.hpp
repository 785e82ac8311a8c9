// Internal interfaces of the MI355X-native TDSE-Z propagation path: host set-up
// (basis, bands, hierarchy) and the launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace tdz {

// ------------------------------------------------------------ host set-up
// One axis of one multigrid level: knots, Gauss points, basis tables and the
// 1D Gram bands (entry (i, j) at i*(2p+1) + j-i+p, as src/tensorops.cpp:27-61).
struct Axis1D {
    int E = 1, p = 0, Q = 1, n = 1;
    bool active = false, wall = false;
    double lo = 0, hi = 0;
    std::vector<double> t;               // knots
    std::vector<double> X, W, cap;       // [E*Q] points, weights, CAP profile
    std::vector<int> off;                // [E] first function of each element
    std::vector<double> B0, B1;          // [(e*(p+1)+a)*Q+q] values, first derivatives
    std::vector<double> M, S, Xb, Wc;    // bands: mass, stiffness, position, CAP profile
    std::vector<double> Vs;              // band of the separable part of s V
    std::vector<char> isWall;            // [n]
};

struct PhysParams {
    double mass = 1, s = 1;              // carrier mass and sign s = -charge
    int pot_kind = 0;
    double pot[4] = {0, 0, 0, 0};
    int cap_kind = 0;
    double cap[4] = {0, 0, 0, 0};
};

// Gauss-Legendre nodes/weights on [-1, 1] (Golub-Welsch).
void gauss_legendre(int Q, double *x, double *w);
// Build one axis from a knot vector (degree p, Q points per non-empty span).
void build_axis(Axis1D &A, const std::vector<double> &knots, int p, int Q, bool wall, const PhysParams &ph);
// Degenerate axis of a lower-dimensional problem: one function equal to 1.
void build_degenerate_axis(Axis1D &A);
// Open uniform knot vector, continuity p-1.
std::vector<double> uniform_knots(int E, double lo, double hi, int p);
// Remove every second simple interior element boundary (the coarsening rule of
// src/multigrid.cpp:25-40); false if nothing was removed.
bool coarsen_knots(const std::vector<double> &U, std::vector<double> &C);
int knot_spans(const std::vector<double> &U);
// Knot-insertion matrix of S_p(C) in S_p(F) as sparse rows (coarse index, weight), fine x
// coarse (Oslo algorithm).
std::vector<std::vector<std::pair<int, double>>> refinement_matrix(const std::vector<double> &C,
                                                                   const std::vector<double> &F, int p);
// max |N_c,j(x) - sum_i R(i, j) N_f,i(x)| over sample points of every fine element
double refinement_error(const std::vector<double> &C, const std::vector<double> &F, int p,
                        const std::vector<std::vector<std::pair<int, double>>> &R);
// the p+1 B-splines of degree p on U nonzero at x (first: index of the first)
void basis_values(const std::vector<double> &U, int p, double x, int &first, double *vals);
std::vector<double> greville(const std::vector<double> &U, int p);
// Symmetric eigen-decomposition (Householder tridiagonalisation + implicit QL):
// A (n x n row-major) is replaced by the eigenvectors as columns.
void sym_eig(int n, std::vector<double> &A, std::vector<double> &lam);
// Generalised T v = lambda M v with M SPD (bands of axis A, restricted to non-wall
// functions): V (n x n, columns = M-orthonormal eigenvectors, zero at walls), lam.
void gen_eig_bands(const Axis1D &A, const std::vector<double> &Tband, std::vector<double> &V,
                   std::vector<double> &lam);

// ------------------------------------------------------------ device side
template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R re, im;
};
using dcomplex = cpx<double>;
using fcomplex = cpx<float>;

// Device copy of the per-axis tables the kernels read.
template <typename R>
struct AxisDevT {
    int E, p, Q, n;
    const R *B0;        // [E][p+1][Q]
    const R *X;         // [E*Q]
    const R *W;         // [E*Q]
    const int *off;     // [E]
};
using AxisDev = AxisDevT<double>;

struct PotParams {
    int kind;
    double s;           // carrier sign
    double a0, a1;      // harmonic: k; soft-core: Z, a
};

// y += g * (s V) x over all elements (generic sum-factorised kernel).
template <typename R>
void launch_pot_apply_generic(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                              cpx<R> *y);
// Same for dim 3 with simple interior knots (tiled z-marching kernel); returns
// false when no instantiation covers (p, Q) or variant. scr: pot3d_scratch_len(N)
// values of scratch (the tiles' contributions by offset class, summed into y in a fixed
// order: the result does not depend on the schedule).
template <typename R>
bool launch_pot_apply_3d(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                         cpx<R> *y, int variant, cpx<R> *scr);
long pot3d_scratch_len(long N);
// number of tile-shape variants of the tiled kernel for (p, Q) (0: none)
int pot3d_variants(int p, int Q);
// 2D grids with simple knots, p in 3..5, Q = 8: tiled kernel, variant = tile shape
// (0: 1x1, 1: 2x1, 2: 2x2, 3: 4x2, 4: 4x4 elements); false if not applicable.
template <typename R>
bool launch_pot_apply_2d(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                         cpx<R> *y, int variant);
int pot2d_variants(int p, int Q);
// Separable Kronecker apply y = Mz My Bx x + Mz By Mx x + Bz My Mx x, bands on device:
// M_d real [n_d*(2p+1)], B_d complex [n_d*(2p+1)]; tmp = 4 scratch vectors.
template <typename R>
struct SepBandsT {
    int n[3], p[3];
    const R *M[3];
    const cpx<R> *B[3];
    // layered z axis: the z factor of the x and y terms (the 1/m(z)-weighted mass band);
    // null: M[2]
    const R *Mz2 = nullptr;
};
using SepBands = SepBandsT<double>;
template <typename R>
void launch_sep_apply(hipStream_t st, const SepBandsT<R> &sb, const cpx<R> *x, cpx<R> *y, cpx<R> *tmp[4]);
// y[I] = x[I] where the wall mask flags I (1D flags per axis on device).
template <typename R>
void launch_wall_fix(hipStream_t st, const int n[3], const unsigned char *wall[3], const cpx<R> *x, cpx<R> *y);

// Deterministic two-stage reductions. partial must hold kMaxBlocks*(nv) entries.
constexpr int kRedBlocks = 512;
// h[i] = <V_i, w>, i < nv (V_i = V + i*ld), result complex on device out[nv]
void launch_multi_dot(hipStream_t st, const dcomplex *V, long ld, int nv, const dcomplex *w, long N,
                      dcomplex *partial, dcomplex *out);
// w -= sum_i h_i V_i (h on device), and out_norm2 = ||w||^2 (on device, deterministic)
void launch_multi_axpy_norm(hipStream_t st, const dcomplex *V, long ld, int nv, const dcomplex *h, dcomplex *w,
                            long N, double *partial, double *out_norm2);
// Classical Gram-Schmidt with the second pass deferred: w -= sum_i h_i V_i (nv <=
// kCgsFusedMax, h on device), out_norm2 = ||w||^2 and out_dot[i] = <V_i, w> (the second
// pass's projections) from one pass; launch_correct_scale then applies w = s (w - V out_dot)
constexpr int kCgsFusedMax = 4;
void launch_multi_axpy_norm_dot(hipStream_t st, const dcomplex *V, long ld, int nv, const dcomplex *h, dcomplex *w,
                                long N, double *pnorm, dcomplex *pdot, double *out_norm2, dcomplex *out_dot);
void launch_correct_scale(hipStream_t st, const dcomplex *V, long ld, int nv, const dcomplex *h, dcomplex *w, long N,
                          double s);
// x = base + sum_i c_i Z_i (c on device; base null: x itself)
// up to kCoefArgMax coefficients passed by value in the kernel arguments (no upload)
constexpr int kCoefArgMax = 64;
struct CoefArg {
    dcomplex v[kCoefArgMax];
};
void launch_multi_axpy(hipStream_t st, const dcomplex *Z, long ld, int nv, const dcomplex *c, dcomplex *x, long N,
                       const dcomplex *base = nullptr);
// the same with the coefficients (host memory, nv <= kCoefArgMax) in the kernel's arguments
void launch_multi_axpy_args(hipStream_t st, const dcomplex *Z, long ld, int nv, const dcomplex *hc, dcomplex *x,
                            long N, const dcomplex *base = nullptr);
// The bands of the shifted operator alpha M + g H (non-layered), from the device copies
// of the axis bands, per axis d (active: d < dim):
//   B_d = g (S_d / 2m - i W_d + sE_d X_d) (+ alpha M_d for d = 0; g (0) on inactive axes),
//   Bc_d = g (-i W_d + sE_d X_d) as complex64 (the separable field part, if Bc_d),
//   Bf_d = B_d as complex64 (if Bf_d).
// W_d is the CAP band (its term only with cap). Evaluated in the same order as the host
// formula, without contraction, so the bands are the same to the bit.
struct BandUpdate {
    const double *M[3] = {}, *S[3] = {}, *W[3] = {}, *X[3] = {};
    dcomplex *B[3] = {};
    fcomplex *Bc[3] = {}, *Bf[3] = {};
    int sz[3] = {0, 0, 0}, active[3] = {0, 0, 0};
    dcomplex alpha{0, 0}, g{0, 0};
    double inv2m = 0, sE[3] = {0, 0, 0};
    int cap = 0;
};
void launch_band_update(hipStream_t st, const BandUpdate &bu);
// y = a x + b y (scalars on host; b = 0: y is not read)
void launch_axpby(hipStream_t st, dcomplex a, const dcomplex *x, dcomplex b, dcomplex *y, long N);
// y = sum_k c_k v_k (k <= kMax terms; y may be one of the v_k: every element is read
// before it is written)
struct LinComb {
    static constexpr int kMax = 24;
    const dcomplex *v[kMax];
    dcomplex c[kMax];
    int k = 0;
    void add(dcomplex coef, const dcomplex *vec)
    {
        if (k >= kMax) throw std::runtime_error("LinComb: more than kMax terms");
        c[k] = coef;
        v[k++] = vec;
    }
};
void launch_lincomb(hipStream_t st, const LinComb &lc, dcomplex *y, long N);
// y1 = sum l1; y2 = y1 + sum l2; y3 = sum l3, one pass
void launch_lincomb3(hipStream_t st, const LinComb &l1, const LinComb &l2, const LinComb &l3, dcomplex *y1,
                     dcomplex *y2, dcomplex *y3, long N);
// the same, with out[0] = ||y2||^2 and out[1] = ||y1||^2 (on device, deterministic;
// partial: scratch of 2 nbmax doubles)
void launch_lincomb3_norms(hipStream_t st, const LinComb &l1, const LinComb &l2, const LinComb &l3, dcomplex *y1,
                           dcomplex *y2, dcomplex *y3, long N, double *partial, int nbmax, double *out);
// m = Mz My Mx x and dz = Xz My Mx x (FP64, real bands of order p on every axis, 3D), and
// xc = x if xc; false if not applicable
bool launch_mass_dipole_z(hipStream_t st, const int n[3], int p, const double *M[3], const double *Xz,
                          const dcomplex *x, dcomplex *m, dcomplex *dz, dcomplex *xc = nullptr);
template <typename R>
void launch_axpby_t(hipStream_t st, cpx<R> a, const cpx<R> *x, cpx<R> b, cpx<R> *y, long N);
// precision conversions
void launch_demote(hipStream_t st, const dcomplex *x, fcomplex *y, long N);
void launch_promote(hipStream_t st, const fcomplex *x, dcomplex *y, long N);
void launch_promote_add(hipStream_t st, const fcomplex *x, const fcomplex *y, dcomplex *z, long N);   // z = x + y
// q = r - g Vg .* My - c - n (n may be null), complex64; with no (the network's [2N]
// output as two real halves) also - ng (no_re + i no_im)
void launch_corr_combine(hipStream_t st, const fcomplex *r, const fcomplex *My, const fcomplex *c,
                         const fcomplex *n, const float *Vg, fcomplex g, fcomplex *q, long N,
                         const float *no = nullptr, fcomplex ng = {0.f, 0.f});
// y = x * s where s = 1/sqrt(*norm2) read on device
void launch_scale_invsqrt(hipStream_t st, const dcomplex *x, const double *norm2, dcomplex *y, long N);
// out = ||x||^2 on device
void launch_norm2(hipStream_t st, const dcomplex *x, long N, double *partial, double *out);

// Dense axis transform y = T x along axis d (T real n x n, row-major; trans: T^T).
// Tfrag: for R = float, the matrix in the fragment order of axis_fragments (the
// register-resident MFMA transform; null: the LDS-tiled kernel); ignored for double
template <typename R>
void launch_dense_axis(hipStream_t st, int d, const int n[3], const R *T, bool trans, const cpx<R> *x, cpx<R> *y,
                       const R *Tfrag = nullptr);
// k-steps of 4 of the register-resident transform for an axis of n (0: axis too long)
int axis_frag_steps(int n);
// n x n row-major T (or its transpose) as the transform's A fragments (empty: too long)
std::vector<float> axis_fragments(int n, const std::vector<double> &T, bool trans);
// Split-FP16 MFMA form of the complex64 transform (each product as hi*hi + hi*lo + lo*hi of
// FP16 halves, FP32 accumulation; about 2^-22 relative): the matrix pre-split and scaled by
// 2^-e (|A| < 1; ainv = 2^e undoes it) in the fragment order of axis_split16_kernel (empty:
// axis too long); false if not applicable.
// FP64 MFMA form (axis_wave_f64_kernel): k-steps of 4 (even, 0 for axes above 160) and the
// matrix (trans: its transpose) in that kernel's fragment order
int axis_f64_steps(int n);
std::vector<double> axis_fragments_f64(int n, const std::vector<double> &T, bool trans);
int axis_split16_blocks(int n);
bool split16_fits(const int n[3]);   // the vectors fit the 32-bit byte offsets of its buffer accesses
std::vector<unsigned short> axis_fragments_split16(int n, const std::vector<double> &T, bool trans, float &ainv);
// split16: 1 a wave per (column tile, three row tiles); 2 a wave per column tile and every
// row tile, the matrix staged in LDS per block
bool launch_dense_axis_split16(hipStream_t st, int d, const int n[3], const unsigned short *Th, float ainv,
                               const fcomplex *x, fcomplex *y, int split16);
bool launch_dense_axis_diag_split16(hipStream_t st, const int n[3], const unsigned short *Th, float ainv,
                                    const float *lam[3], const unsigned char *wall[3], fcomplex alpha, fcomplex g,
                                    const fcomplex *x, fcomplex *y, int split16);
// One thread per output, T read through the caches (reference kernel for tests).
// complex64 forward transform along z fused with the diagonal solve (Tfrag: fragments
// of T^T; then divide by alpha + g (l0 + l1 + l2) off the walls); false if not applicable
bool launch_dense_axis_diag(hipStream_t st, const int n[3], const float *Tfrag, const float *lam[3],
                            const unsigned char *wall[3], fcomplex alpha, fcomplex g, const fcomplex *x, fcomplex *y);
template <typename R>
void launch_dense_axis_simple(hipStream_t st, int d, const int n[3], const R *T, bool trans, const cpx<R> *x,
                              cpx<R> *y);
// y[i,j,k] = x[i,j,k] / (alpha + g (lx[i] + ly[j] + lz[k])), walls: y = x
template <typename R>
void launch_diag_scale(hipStream_t st, const int n[3], const R *lam[3], const unsigned char *wall[3], cpx<R> alpha,
                       cpx<R> g, const cpx<R> *x, cpx<R> *y);

// Jacobi-Chebyshev smoother: dinv = 1 / diag(A) of a level (Kronecker bands, plus g times
// the potential diagonal from the Galerkin stencil Vc or the lumped values Vl times diag M;
// walls: 1), and the recurrence update d = a d + b dinv .* res, e += d (e may be null).
template <typename R>
void launch_jacobi_diag(hipStream_t st, const int n[3], const int p[3], const R *M[3], const cpx<R> *B[3],
                        const R *Vc, const R *Vl, cpx<R> g, const unsigned char *wall[3], cpx<R> *dinv);
template <typename R>
void launch_cheb_update(hipStream_t st, long N, R a, R b, const cpx<R> *dinv, const cpx<R> *res, cpx<R> *d,
                        cpx<R> *e);

// Sparse 1D transfer along axis d: out (axis d of size m) = A in (axis d of size nin),
// A in CSR (rowptr[m+1], col, val); other axes unchanged. accumulate: out += A in.
template <typename R>
void launch_csr_axis(hipStream_t st, int d, const int nin[3], int m, const int *rowptr, const int *col,
                     const R *val, const cpx<R> *in, cpx<R> *out, bool accumulate);

// x = A_ij^-1 rhs along z for every (i, j) line, A_ij = Bz + g (lx[i] + ly[j]) Mw (banded,
// half-width p; banded elimination without pivoting, one thread per line; work holds
// n0 n1 n2 (2p+1) values). The separable solve of a layered structure in the x/y
// eigenbasis.
template <typename R>
void launch_zline_solve(hipStream_t st, const int n[3], int p, const R *lx, const R *ly, const cpx<R> *Bz,
                        const R *Mw, cpx<R> g, const cpx<R> *rhs, cpx<R> *x, cpx<R> *work);

// ---------------------------------------------------------------- allocation-free entry
// Reductions of tdz_observables_noalloc: kObsDots dot products of kRedBlocks partials
// each (one set at a time) and their results.
constexpr int kObsDots = 8;
constexpr long noalloc_scratch_len() { return (long)kObsDots * kRedBlocks + kObsDots; }
// resolve the device address of the static scratch on the current device (once per device)
void noalloc_init();

// ---------------------------------------------------------------- exact Galerkin potential
// Per axis: for every (function i, offset d) pair of a coarse level (index (d+p) n + i),
// the products N_i N_{i+d} at the run of the finest level's Gauss points [q0, q0+len)
// where both are nonzero (values from val + off).
struct GalPairs {
    int npairs = 0;
    const int *q0 = nullptr, *len = nullptr, *off = nullptr;
    const double *val = nullptr;
};
struct GalerkinPlan {
    GalPairs pairs[3];
    const double *X[3] = {}, *W[3] = {};   // finest level's Gauss points and weights per axis
    int nq[3] = {1, 1, 1};
    PotParams pot{};
};
// S = P^T (s V) P of one coarse level as an offset-major stencil (zeroed first by the
// first slab); work: work_len doubles of scratch
void galerkin_potential(hipStream_t st, const GalerkinPlan &g, double *S, double *work, long work_len);
void launch_stencil_demote(hipStream_t st, const double *S, float *Sf, long n);
// y += g S x for an offset-major stencil of half-widths p[a] (0 on unused axes)
template <typename R>
void launch_stencil_apply(hipStream_t st, const int n[3], const int p[3], const R *S, R gr, R gi, const cpx<R> *x,
                          cpx<R> *y);

// ---------------------------------------------------------------- device FGMRES(k)
constexpr int kGmMax = 8;   // longest captured cycle
struct GmState {
    dcomplex H[(kGmMax + 1) * kGmMax];
    double cs[kGmMax];
    dcomplex sn[kGmMax];
    dcomplex gv[kGmMax + 1];
    double bnorm;
    int done;
};
void launch_gm_start(hipStream_t st, const double *beta2, const double *bnorm2, GmState *g);
// column j of an m-column cycle from h = h1 + h2 (the two CGS passes) and ||w||^2
void launch_gm_arnoldi(hipStream_t st, int j, int m, const dcomplex *h1, const dcomplex *h2, const double *wn2,
                       GmState *g);
void launch_gm_solve(hipStream_t st, int k, int m, GmState *g, dcomplex *y, double *res, const int *step);
void launch_step_advance(hipStream_t st, int *step);
// out = binomial extrapolation of order min(order, *step) from x, h1, h2 (device step)
void launch_extrapolate(hipStream_t st, const int *step, int order, const dcomplex *x, const dcomplex *h1,
                        const dcomplex *h2, dcomplex *out, long N);
void launch_field_band(hipStream_t st, long n, const double *M, const dcomplex *T0, const double *X, dcomplex g,
                       const double *sE, const int *step, int d, dcomplex *B);

// ---------------------------------------------------------------- learned part
constexpr int kCnnChannels = 16;
constexpr int kCnnInputs = 4;   // y and its three first-moment features
// FP32 forward: m [2N][C] work, out [2N] (re then im component)
void cnn_forward_f32(hipStream_t st, const int n[3], const fcomplex *u, const float *W1, const float *W2,
                     const float *P, float *m, float *out);
// FP32 backward of the loss sum (out - t)^2 * inv_t2 over both components
void cnn_backward_f32(hipStream_t st, const int n[3], const fcomplex *u, const fcomplex *target, float inv_t2,
                      const float *W2, const float *P, const float *m, const float *out, float *dov, float *dh,
                      float *partial, int nchunk, float *dW1, float *dW2);
// The positional fields as closed forms of the Greville coordinates (x, y, z) =
// (gr[0][i], gr[1][j], gr[2][k]): field q of voxel (i, j, k) is scale[q] f_q with
// rho = (r^2 + a)^-1/2, V = -sZ rho,
//   f_0..3   = 1, V, sZ x rho^3 h_x, sZ y rho^3 h_y
//   f_4..7   = sZ z rho^3 h_z, sZ (rho^3 - 3 x_d^2 rho^5) h_d^2 (d = x, y, z)
//   f_8..11  = -3 sZ x_d x_e rho^5 h_d h_e (xy, xz, yz), V^2
//   f_12..15 = x / L_x, y / L_y, z / L_z, r^2 inv_r2
// (the host's FP32 table [N][C] of the same fields feeds training).
struct CnnFieldGen {
    const float *gr[3] = {nullptr, nullptr, nullptr};
    float sZ = 0.f, a = 1.f, h[3] = {0.f, 0.f, 0.f}, inv_l[3] = {0.f, 0.f, 0.f}, inv_r2 = 0.f;
    float scale[kCnnChannels] = {};
};
// BF16 MFMA forward; mb work of cnn_infer_work_bytes(N) bytes ([2N][C] bf16 plus the
// packed weights), out [2N]; if nout: nout = g (out_re + i out_im). The positional fields
// are evaluated in FP32 from fg where they are used. pack: pack W1, W2 into their BF16
// operand order in mb first (else mb holds them from an earlier call).
inline size_t cnn_infer_work_bytes(long N) { return (size_t)2 * N * kCnnChannels * 2 + 64 * 64 * 8; }
void cnn_infer_mfma(hipStream_t st, const int n[3], const fcomplex *u, const float *W1, const float *W2,
                    const CnnFieldGen &fg, unsigned short *mb, float *out, fcomplex g, fcomplex *nout,
                    bool pack = true);
void cnn_to_complex(hipStream_t st, long N, const float *out, fcomplex g, fcomplex *n);
// partial[b] (b < kLossBlocks, FP64) sums to sum (out - t)^2 over both components
constexpr int kLossBlocks = 256;
void cnn_loss(hipStream_t st, long N, const float *out, const fcomplex *target, double *partial);
// r = complex values with both parts uniform in [-1, 1), reproducible from seed
void random_field(hipStream_t st, long N, unsigned long long seed, fcomplex *r);
// t = pv - Vg .* My
void lumped_residual(hipStream_t st, long N, const fcomplex *pv, const fcomplex *My, const float *Vg, fcomplex *t);
// u[c] *= v[c], c < kCnnInputs (in place)
struct CnnScales { float v[kCnnInputs]; };
void scale_inputs(hipStream_t st, long N, const CnnScales &s, fcomplex *u);
// Fused products of the network-corrected solve (complex64): My, corr (the C bands'
// Kronecker sum applied to y) and, if u is not null, the scaled network input
// u = [s0 y, s1 G_x, s2 G_y, s3 G_z], G_d = X_d y - x_d My. tmp: 7 work vectors.
struct FusedBands {
    int n[3], p[3];
    const float *M[3], *X[3];    // real mass and position bands
    const fcomplex *C[3];        // complex bands of the C term
    const float *g[3];           // Greville coordinates
};
// 3D with equal orders: the x/y products of every plane into tmp[0..3], then the z passes;
// otherwise three separate passes.
void fused_products(hipStream_t st, const FusedBands &fb, const CnnScales &sc, const fcomplex *y, fcomplex *tmp[7],
                    fcomplex *My, fcomplex *corr, fcomplex *u);
// The right-hand side of a correction sweep without the network, q = r - g Vg (M y) - corr(y)
// (corr: the CAP and field bands; r null: zero), in two passes: the x and y band passes of
// every plane into s1, s2 (complex64 scratch vectors), then the z band passes with the
// combine. False where it does not apply (unequal orders, not 3D): the caller then takes
// fused_products and the combine.
bool fused_correction(hipStream_t st, const FusedBands &fb, const fcomplex *y, const fcomplex *r, const float *Vg,
                      fcomplex g, fcomplex *q, fcomplex *s1, fcomplex *s2);

}  // namespace tdz
