/*
 * tdz.h — C ABI of the MI355X-native TDSE-Z propagation hot path.
 *
 * The reference solves every Crank-Nicolson step through PETSc: TSSetIJacobian /
 * TSSetIFunction callbacks build J = i a M - H - s E(t) D (src/propagator.cpp:103-105,
 * src/monitor.cpp:150-162, :367-382), KSPSolve runs (F)GMRES with a right
 * preconditioner (src/propagator.cpp:131-197) and PCMG applies the geometric
 * B-spline multigrid (src/multigrid.cpp:205-327). With PETSc present, the
 * plug-in points are a MATSHELL MATOP_MULT and a PCSHELL apply; the entry points
 * below have exactly those shapes, on device-resident interleaved complex128
 * vectors (re, im pairs, x fastest: index i + n0 (j + n1 k), as
 * src/tensorops.cpp:67-70).
 *
 * Conventions (mirroring PetscErrorCode): every function returns 0 on success and
 * a nonzero code on failure; tdz_last_error() describes the last failure of the
 * calling thread. A solve that does not converge is an error
 * (src/propagator.cpp:671-672). The creator destroys: the context owns all
 * device memory it allocates, the caller owns host buffers. One host thread and
 * one HIP stream per context.
 */
#ifndef TDZ_H
#define TDZ_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdz_ctx tdz_ctx;

/* Potential kinds: s V(r) enters H (s = -charge, include/tdsez_parser.hpp:85-105). */
enum { TDZ_POT_NONE = 0, TDZ_POT_HARMONIC = 1 /* 0.5 pot[0] r^2 */, TDZ_POT_SOFTCORE = 2 /* -pot[0]/sqrt(r^2+pot[1]) */ };
/* CAP kinds: CAP = -i sum_d W_d(x_d) (src/assembly/cap.cpp:1-21). */
enum { TDZ_CAP_NONE = 0, TDZ_CAP_QUADRATIC = 1 /* cap[0] (cap[1]-s)^2 within cap[1] of a wall */,
       TDZ_CAP_MANOLOPOULOS = 2 /* kmin = cap[0] (src/assembly/assembly.cpp:170-208) */ };
/* Preconditioners. */
enum { TDZ_PC_NONE = 0, TDZ_PC_SEPARABLE = 1 /* fast-diagonalisation inverse of alpha M + g (K + V_sep) */,
       TDZ_PC_MG = 2 /* geometric B-spline multigrid V-cycle */,
       TDZ_PC_SEPARABLE_MIXED = 3 /* the separable solve in complex64 */,
       TDZ_PC_MG_MIXED = 4 /* the V-cycle in complex64: smoother, level operators, transfers */,
       TDZ_PC_CORRECTED = 5 /* complex64 first-order corrected separable solve: S^-1 (r - dA S^-1 r), dA = the
                               CAP and field terms exactly plus the potential lumped at the Greville points */,
       TDZ_PC_CORRECTED_EXACT = 6 /* the same with the potential product by the complex64 matrix-free kernel */,
       TDZ_PC_LEARNED = 7 /* TDZ_PC_CORRECTED plus the learned remainder of the potential product (BF16 MFMA
                             CNN); falls back to plain multigrid when Krylov iterations regress */ };
/* Single operator terms for tdz_term_apply. */
enum { TDZ_TERM_M = 0, TDZ_TERM_K = 1, TDZ_TERM_V = 2, TDZ_TERM_CAP = 3, TDZ_TERM_DX = 4, TDZ_TERM_DY = 5,
       TDZ_TERM_DZ = 6 };

/* Grid and physics of one run: the subset of the reference input deck that the
 * propagation path reads (include/tdsez_parser.hpp:270-542). */
typedef struct {
    int dim;             /* 1..3 */
    int nelem[3];        /* uniform elements per axis (reference: Nelements, one value) */
    double lo[3], hi[3]; /* domain per axis */
    int p;               /* SplineDegree, continuity p-1 (src/core.cpp:1899) */
    int nquad;           /* NQuadratures, Gauss-Legendre (default 8) */
    int dirichlet;       /* 1: wall DOFs at both ends of every active axis */
    double mass;         /* constant Mass */
    double charge;       /* Charge (-1 electron) */
    int pot_kind;
    double pot[4];
    int cap_kind;
    double cap[4];
    int mg_levels;       /* MGLevels: 0 automatic, else the number of levels */
    int mg_smooth;       /* MGSmooth: smoothing sweeps per level (default 1) */
    int restart;         /* Krylov restart; 0 = reference rule (30, 50 above 1e5 rows) */
    /* z axis of a layered structure (dim 3). knots: an open knot vector for z (nknots
     * values; interior knots repeated p times give C0 interfaces), or null for uniform.
     * nlayers > 0: piecewise-constant layers [layer_bounds[i], layer_bounds[i+1]) of
     * effective mass layer_mass[i] and potential offset layer_v[i] (enters as s V):
     * BenDaniel-Duke kinetic term 1/2 grad.(1/m(z)) grad. With layers the CAP, fields
     * along x or y and the harmonic potential are not supported. The arrays are read
     * by tdz_create only. */
    const double *knots;
    int nknots;
    int nlayers;
    const double *layer_bounds;
    const double *layer_mass;
    const double *layer_v;
} tdz_problem;

typedef struct {
    int its;          /* Krylov iterations of the last solve */
    int restarts;     /* restart cycles */
    double res;       /* final relative residual ||b - A x|| / ||b|| (estimate of right-preconditioned GMRES) */
    double bnorm;     /* ||b|| */
    int pc_rebuilds;  /* preconditioner refreshes (TDSEZPCRefresh) */
} tdz_ksp_stats;

/* sin^2-envelope pulse E_a(t) = E0[a] sin^2(pi t / T) sin(omega t + cep), 0 <= t <= T,
 * T = ncycles 2 pi / omega; zero outside. */
typedef struct {
    double E0[3];
    double omega;
    double ncycles;
    double cep;
} tdz_pulse;

/* Lifetime. Replaces the IGA/assembly set-up that feeds the propagator
 * (src/core.cpp:1786-2081, src/tensorops.cpp:16-81, src/assembly/cap.cpp:40-144). */
int tdz_create(const tdz_problem *prob, tdz_ctx **ctx);
int tdz_destroy(tdz_ctx *ctx);
const char *tdz_last_error(void);
long tdz_size(const tdz_ctx *ctx);
int tdz_dims(const tdz_ctx *ctx, int n[3]);
int tdz_levels(const tdz_ctx *ctx);

/* Device vectors of N complex128 owned by the context (VecCreate/VecDestroy). */
int tdz_vec_alloc(tdz_ctx *ctx, double **dvec);
int tdz_vec_free(tdz_ctx *ctx, double *dvec);
int tdz_vec_upload(tdz_ctx *ctx, double *dvec, const double *host);
int tdz_vec_download(tdz_ctx *ctx, const double *dvec, double *host);
int tdz_sync(tdz_ctx *ctx);

/* A = alpha M + g (H + sum_a sE_a D_a), H = K + s V + CAP: what the IJacobian
 * callback would record into a shell context (src/monitor.cpp:367-382 with
 * alpha = i a, g = -1 there; CN uses alpha = 1, g = i dt/2). */
int tdz_set_shift(tdz_ctx *ctx, double alpha_re, double alpha_im, double g_re, double g_im, const double sE[3]);
/* MATSHELL MATOP_MULT: y = A x (replaces MatMult(J, x, y) inside KSPSolve). */
int tdz_op_apply(tdz_ctx *ctx, const double *x, double *y);
/* y = T x for one term of H, M or D (TDSEZTensorOps::Apply, src/tensorops.cpp:155-188). */
int tdz_term_apply(tdz_ctx *ctx, int term, const double *x, double *y);
/* PCSHELL set-up and apply (replaces PCSetUp/PCApply of PCMG, src/multigrid.cpp:205-327). */
int tdz_pc_setup(tdz_ctx *ctx, int kind);
int tdz_pc_apply(tdz_ctx *ctx, const double *r, double *z);
/* KSPSolve: FGMRES, right preconditioning, x holds the initial guess
 * (src/propagator.cpp:131-182). Not converging within maxit is an error. */
int tdz_ksp_solve(tdz_ctx *ctx, const double *b, double *x, double rtol, int maxit, tdz_ksp_stats *stats);
/* TSSetIFunction callback body (TDSEZIFunctionPol{X,Y,Z}, src/monitor.cpp:108-162):
 * F = i M psidot - H psi - sum_a sE_a D_a psi with sE = s E(t) (H includes the CAP);
 * F = 0 on Dirichlet wall DOFs. The shift set by tdz_set_shift is left unchanged. */
int tdz_ifunction(tdz_ctx *ctx, const double sE[3], const double *psi, const double *psidot, double *F);

/* Multigrid levels (level 0 the finest; the hierarchy of src/multigrid.cpp:239-268).
 * tdz_mg_transfer: dir 0 restricts out = P^T in (level -> level+1), dir 1 prolongs
 * out = P in (level+1 -> level), P the knot-insertion prolongation (the interpolation
 * PCMGSetInterpolation receives, src/multigrid.cpp:295-298; restriction P^T). Device
 * vectors of at least the level sizes (tdz_level_dims).
 * tdz_level_apply: y = A_l x with the current shift: A_0 = tdz_op_apply's operator,
 * coarse levels P^T A P (PCMGSetGalerkin(PC_MG_GALERKIN_BOTH), src/multigrid.cpp:277). */
int tdz_level_dims(const tdz_ctx *ctx, int level, int n[3]);
int tdz_mg_transfer(tdz_ctx *ctx, int level, int dir, const double *in, double *out);
int tdz_level_apply(tdz_ctx *ctx, int level, const double *x, double *y);
/* -tdsez_pc_refresh_its N (include/tdsez_propagator.hpp:22-24): besides the rule
 * its > 2 base + 5, rebuild the preconditioner whenever a solve takes more than
 * max_its iterations (max_its < 0: off, the default). Rebuilds are counted in
 * tdz_ksp_stats.pc_rebuilds. */
int tdz_set_pc_refresh(tdz_ctx *ctx, int max_its);
/* Multigrid smoother on every level but the coarsest (which keeps its separable sweeps):
 * kind 0 (default) Richardson with the level's fast-diagonalisation separable solve,
 * kind 1 Jacobi-Chebyshev of the given degree (degree - 1 level applies per sweep) over
 * [lmax/30, lmax] of D^-1 M, D = diag of the level operator (the mixed-precision
 * configuration's "FP32 Chebyshev smoother" under TDZ_PC_MG_MIXED; the reference smooths
 * with Richardson + restricted ASM/ILU(0), src/multigrid.cpp:303-313). */
int tdz_mg_set_smoother(tdz_ctx *ctx, int kind, int degree);
/* Coarsest-level solve of the multigrid preconditioners (the reference: KSPPREONLY with an
 * LU factorisation, PCLU/MUMPS, src/multigrid.cpp:314-325): kind 0 (default) the separable
 * solve followed by MGSmooth Richardson sweeps with it; kind 1 a Krylov solve to the relative residual rtol
 * (GMRES with the separable solve as right preconditioner, restarted every min(8, maxit)
 * iterations, at most maxit iterations rounded up to whole cycles, in FP64 also under
 * TDZ_PC_MG_MIXED). tdz_mg_coarse_status: the iterations
 * and residual estimate of the last Krylov coarse solve. */
int tdz_mg_set_coarse(tdz_ctx *ctx, int kind, double rtol, int maxit);
int tdz_mg_coarse_status(const tdz_ctx *ctx, int *its, double *res);
/* x = the coarse solve of the multigrid preconditioners applied to b on the coarsest level
 * (level = tdz_levels - 1 >= 1; PCMGGetCoarseSolve + KSPSolve on it), with the current
 * shift and the kind tdz_mg_set_coarse chose. Device vectors of the level's size. */
int tdz_level_solve(tdz_ctx *ctx, int level, const double *b, double *x);
/* Transforms of the complex64 separable solves (every mixed-precision preconditioner; no
 * reference counterpart): kind 0 FP32 MFMA, kind 1 (default) split FP16 MFMA (each operand
 * the sum of two FP16 halves, three FP16 products in FP32 accumulation, about 2^-22 relative
 * per product; axes up to 160 functions, longer ones keep FP32), kind 2 the same products
 * with the matrix staged in LDS per block (a wave per column tile and every row tile),
 * kind 3 single FP16 products (about 2^-11 relative; what the correction form below uses
 * for its corrections). */
int tdz_set_fd_transform(tdz_ctx *ctx, int kind);
/* Correction sweeps of TDZ_PC_CORRECTED* and TDZ_PC_LEARNED with the split-FP16 transforms (no
 * reference counterpart): form 1 (default) keeps y = S^-1 r and solves each sweep for its
 * correction alone, z = y + S^-1 (-dA~ z), on single FP16 products (the correction is about
 * 1e-2 of z, so about 1e-5 of z in rounding); form 0 solves z = S^-1 (r - dA~ z) whole. */
int tdz_set_correction_form(tdz_ctx *ctx, int form);
/* Transforms of the FP64 separable solves (TDZ_PC_SEPARABLE and the FP64 multigrid levels;
 * no reference counterpart): kind 1 (default) FP64 MFMA for axes up to 160 functions, kind 0
 * the VALU form (LDS-tiled), which longer axes keep. */
int tdz_set_fd64_transform(tdz_ctx *ctx, int kind);
/* Correction sweeps of TDZ_PC_CORRECTED* and TDZ_PC_LEARNED (no reference counterpart):
 * z <- S^-1 (r - dA~ z) applied sweeps times after z = S^-1 r (1..4, default 1), dA~ the
 * terms the separable solve S leaves out (CAP and field exactly, the potential lumped,
 * plus the network's remainder for TDZ_PC_LEARNED). */
int tdz_pc_set_corrections(tdz_ctx *ctx, int sweeps);
/* Host utilities of the multigrid set-up (no device work): one coarsening of a knot
 * vector U[m] into C (at most m values; *mc = its length, 0 when nothing can be removed;
 * src/multigrid.cpp:25-40), and the dense knot-insertion matrix R [nf][nc] (row-major,
 * nf = mf - p - 1, nc = mc - p - 1) with coarse B-spline j = sum_i R[i][j] fine B-spline i
 * (src/multigrid.cpp:52-81). */
int tdz_coarsen_knots(const double *U, int m, double *C, int *mc);
int tdz_refinement_matrix(const double *C, int mc, const double *F, int mf, int p, double *R);
/* The last warning of the calling thread (also printed to stderr), e.g. the strong-field
 * warning of multigrid runs (tau max|E.r| > 1.5, src/multigrid.cpp:213-237). */
const char *tdz_last_warning(void);
/* One Crank-Nicolson step psi <- (M + i dt/2 H_E)^-1 (M - i dt/2 H_E) psi with
 * H_E at the midpoint field sE_mid (TSTHETA theta = 1/2 in src/propagator.cpp:94-101). */
int tdz_step_cn(tdz_ctx *ctx, double *psi, double dt, const double sE_mid[3], double rtol, int maxit,
                tdz_ksp_stats *stats);
/* nsteps CN steps from t0 under a pulse (TSSolve of TDSEZPropagator::Evolve,
 * src/propagator.cpp:78-251), with the preconditioner refresh rule of
 * TDSEZPCRefresh (include/tdsez_propagator.hpp:17-31). obs (optional, host,
 * 7 doubles per step) receives [t, norm, Re<H>, Im<H>, <x>, <y>, <z>] after each
 * step; its (optional, host) the Krylov iterations per step. */
int tdz_evolve_cn(tdz_ctx *ctx, double *psi, double t0, double dt, int nsteps, const tdz_pulse *pulse,
                  double rtol, int maxit, double *obs, int *its);
/* Initial guess of the CN solves of tdz_evolve_cn: order 0 (default) starts each
 * solve from the current state, as TSTHETA does (src/propagator.cpp:149); order
 * k = 1..4 from the binomial extrapolation of the last k+1 states (the guess of the
 * stage solves, src/propagator.cpp:629-638), whose residual the library forms from
 * stored products instead of an operator apply. Every solve still meets rtol on its
 * residual; the history carries over to the next tdz_evolve_cn call when psi is
 * unchanged. */
int tdz_set_cn_predictor(tdz_ctx *ctx, int order);
/* The predictor's stored products are recomputed exactly every sync_every steps (an
 * operator apply per kept state): carried from solve to solve their rounding would grow
 * like n^(order + 1/2). last_drift is the relative inconsistency found at the last
 * recomputation; the interval adapts to keep it between 1e-3 and 1e-2 of rtol. */
int tdz_cn_predictor_status(const tdz_ctx *ctx, int *sync_every, double *last_drift);
/* Initial states. tdz_project_gaussian: L2 projection of the packet
 * prod_a (2 pi w_a^2)^(-1/4) exp(-(x_a-c_a)^2/(4 w_a^2) + i k_a x_a), normalised so
 * psi^H M psi = 1 (src/core.cpp:1188-1262). tdz_ground_state: lowest eigenstate of
 * H = K + sV (CAP excluded, as before src/propagator.cpp:51-69) by shifted inverse
 * iteration from psi, shift sigma below the ground energy, until the relative
 * eigen-residual < tol (replaces the SLEPc solve of src/core.cpp:758-1083). */
int tdz_project_gaussian(tdz_ctx *ctx, const double centre[3], const double width[3], const double momentum[3],
                         double *psi);
int tdz_ground_state(tdz_ctx *ctx, double *psi, double sigma, double tol, int maxit, double *energy,
                     double *residual, int *iterations);
/* One Pade(2,2) step psi <- R22(-i h M^-1 H_E) psi: stages
 * (M + g_k H_E) y = (M + conj(g_k) H_E) x, g_{1,2} = h(-+sqrt3 + 3i)/12, each
 * started from x plus the order-3 binomial extrapolation of the increments of
 * earlier calls with the same slot (Pade22Step, src/propagator.cpp:564-676). */
int tdz_step_pade22(tdz_ctx *ctx, double *psi, double h, const double sE[3], int slot, double rtol, int maxit,
                    tdz_ksp_stats *stats);
/* nsteps of the rational propagators (EvolveRational, src/propagator.cpp:682-769):
 * magnus = 0: Pade(2,2) with the midpoint field; magnus = 1: commutator-free
 * Magnus 4, two Pade(2,2) half steps with fields 2(a2 E1 + a1 E2), 2(a1 E1 + a2 E2)
 * at the Gauss nodes. obs/its as for tdz_evolve_cn (its: sum over the stages). */
int tdz_evolve_rational(tdz_ctx *ctx, double *psi, double t0, double dt, int nsteps, const tdz_pulse *pulse,
                        int magnus, double rtol, int maxit, double *obs, int *its);
/* Observables [norm psi^H M psi, Re psi^H H psi, Im psi^H H psi, <x>, <y>, <z>]
 * (MonitorProducts, src/manager.cpp:584-672; src/monitor.cpp:1145-1194). */
int tdz_observables(tdz_ctx *ctx, const double *psi, double out[6]);
/* One row of each trace of the observables file for the device state psi at time t
 * (column layouts of src/monitor.cpp:1281-1330, written by tdz_io_trace_append):
 * dip [1 + 3 dim] = [t, E_a, q <D_a>, acceleration_a (0: not computed)],
 * energy [7] = [t, Kin, Pot, Int, Tot, <1/m>, Norm] (real parts; Int = s E.<D>,
 * Tot = <H> + Int), ac [3] = [t, Re, Im psi0^H M psi] (zeros when psi0 is null).
 * E is the field E_a(t) itself (not s E). Any output may be null. */
/* Guess history of the Pade/Magnus stage solves (4 slots x order increments of N complex
 * values, order of tdz_set_pade_guess, default 3; CheckpointGuess,
 * src/propagator.cpp:1546-1590): set = 0 copies the counts and, if hist is not null, the
 * increments to hist[((slot * order) + j) * N] (host); set = 1 restores them. */
int tdz_pade_history(tdz_ctx *ctx, int set, int counts[4], double *hist);
/* -pade_extrapolate_guess and -pade_guess_order (src/propagator.cpp:723-726): whether the
 * stage solves start from the extrapolated increments, and how many (1..5, default 3).
 * The history of tdz_pade_history then holds [4 slots][order][N]. */
int tdz_set_pade_guess(tdz_ctx *ctx, int extrapolate, int order);
int tdz_trace_row(tdz_ctx *ctx, const double *psi, const double *psi0, double t, const double E[3], double *dip,
                  double *energy, double *ac);
/* Allocation-free variant of tdz_observables (no counterpart in the reference; for callers
 * that cannot allocate: code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). The same six values, with the products on the device and the reductions'
 * partial sums in a static device array of the library (65.7 KB per GPU, any N: a fixed
 * number of partials per reduction); no device or host allocation during the call. Calls
 * that overlap on one GPU share that array and race: serialise them. Not for layered
 * structures. */
int tdz_observables_noalloc(tdz_ctx *ctx, const double *psi, double out[6]);
/* Field sE_a(t) = s E_a(t) of a pulse (TDSEZParser::Coupling, include/tdsez_parser.hpp:190-194). */
int tdz_pulse_coupling(const tdz_ctx *ctx, const tdz_pulse *pulse, double t, double sE[3]);

/* Learned correction (no reference counterpart: the reference has no learned
 * component). The network input is u = [y, G_x, G_y, G_z] (4 complex vectors of
 * N), y = S^-1 r the separable solve of the preconditioner's residual and
 * G_d = X_d y - x_d .* M y its first moments about the Greville points; the target
 * is t = V y - Vg .* M y (the part of the potential product the lumped form
 * misses, by the matrix-free kernel). tdz_learn_train harvests (u, t) pairs from
 * the solver's own preconditioned vectors during harvest_steps CN steps from psi
 * (psi is not modified), then fits the network by Adam with the FP32
 * forward/backward kernels; loss per sample |N(u) - t|^2 / |t|^2 (1 = the loss of
 * a zero output), reported as the mean over the samples before and after the fit.
 * It also records the Krylov iterations of the preconditioner without the network
 * (the regression threshold of the fallback).
 * tdz_learn_weights gets (set = 0) or sets (set = 1) W1 [27][4][16], W2 [27][16][16];
 * tdz_learn_fields copies the positional fields [N][16]; tdz_learn_features
 * computes u [4][N] (host, complex interleaved) for a device vector y;
 * tdz_learn_eval runs the FP32 forward/backward and the BF16 MFMA forward on one
 * host pair (u [4][N], t [N], complex interleaved) (host outputs: loss, dW1, dW2,
 * out_f32 and out_mfma [2N] re then im). */
int tdz_learn_train(tdz_ctx *ctx, const double *psi, double t0, double dt, const tdz_pulse *pulse, int harvest_steps,
                    int max_samples, int iters, double lr, double rtol, double *loss_before, double *loss_after);
/* Training set of tdz_learn_train: besides the harvested pairs, random_samples synthetic
 * inputs y = S^-1 w (and S^-1 M S^-1 w) from random fields w drawn from seed, which cover
 * the whole grid (default 6, seed 12345). */
int tdz_learn_config(tdz_ctx *ctx, int random_samples, unsigned long long seed);
int tdz_learn_weights(tdz_ctx *ctx, float *W1, float *W2, int set);
int tdz_learn_fields(tdz_ctx *ctx, float *P);
int tdz_learn_features(tdz_ctx *ctx, const double *y, double *u);
int tdz_learn_eval(tdz_ctx *ctx, const double *u, const double *target, double *loss, float *dW1, float *dW2,
                   double *out_f32, double *out_mfma);
int tdz_learn_status(tdz_ctx *ctx, int *base_its, int *fallbacks, int *active);
/* Choices of the adaptive corrected preconditioners since tdz_pc_setup (no reference
 * counterpart): counts[4 net + k] = solves that ran k correction sweeps (1..3), with
 * (net = 1) or without (net = 0) the network in the last one. */
int tdz_pc_sweep_stats(tdz_ctx *ctx, long *counts);

/* Batched scan (no reference counterpart: the reference runs a scan as separate
 * jobs). nruns independent problems (any grids, degrees and pulses) advance together:
 * one CN step of every run is captured once in a HIP graph (one branch per run: the
 * field bands of the step from a precomputed s E(t_n + dt/2) table, the right-hand
 * side, kiters FGMRES iterations with the Hessenberg problem on the device, the
 * update) and each step is one graph launch. A run whose residual estimate after
 * the captured iterations is above rtol finishes that step on the host FGMRES
 * (counted in tdz_batch_fixups), so every step meets the same FP64 test as
 * tdz_step_cn. pc_kind: TDZ_PC_SEPARABLE, TDZ_PC_SEPARABLE_MIXED or TDZ_PC_NONE;
 * 1 <= kiters <= 8 and <= restart. nsteps bounds the steps the field table covers.
 * tdz_batch_context gives a run's context (for set-up, e.g. tdz_ground_state on psi)
 * and its state vector (device, owned by the batch). its and res ([nsteps][nruns])
 * receive the iterations and relative residuals of each step. */
typedef struct tdz_batch tdz_batch;
int tdz_batch_create(const tdz_problem *probs, const tdz_pulse *pulses, int nruns, double dt, int nsteps,
                     int pc_kind, int kiters, tdz_batch **out);
int tdz_batch_context(tdz_batch *b, int run, tdz_ctx **ctx, double **psi);
int tdz_batch_evolve(tdz_batch *b, int nsteps, double rtol, int maxit, int *its, double *res);
int tdz_batch_fixups(const tdz_batch *b, long *fixups);
/* Initial guess of the captured steps (before the first tdz_batch_evolve): order 1 or 2
 * starts each run's solve from the binomial extrapolation of its last 2 or 3 states
 * (as tdz_set_cn_predictor; its residual then takes one more operator apply in the
 * graph), order 0 (default) from the current state. */
int tdz_batch_set_predictor(tdz_batch *b, int order);
int tdz_batch_destroy(tdz_batch *b);

/* Measurement hooks (no reference counterpart; used by bench.py): HIP events
 * recorded on the context's stream around every launch of the fine-level
 * potential kernel while enabled; stats returns their summed duration and count. */
int tdz_timing(tdz_ctx *ctx, int enable);
int tdz_timing_stats(tdz_ctx *ctx, double *total_ms, long *launches);

#ifdef __cplusplus
}
#endif
#endif /* TDZ_H */
