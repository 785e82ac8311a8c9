// HIP kernels of the TDSE-Z propagation path for gfx950 (CDNA4, wave64).
//
//   pot_apply_generic : matrix-free potential apply y += g (s V) x, one element
//                       per workgroup, sum factorisation in LDS: the forward
//                       evaluation of src/quadobs.cpp:181-238 at the Gauss points,
//                       times g s V(r) w (src/quadobs.cpp:139), and its transpose.
//   sep_*             : separable Kronecker terms by three 1D band passes
//                       (src/tensorops.cpp:97-153), complex bands.
//   multi_dot, ...    : Krylov vector kernels with deterministic two-stage
//                       reductions (fixed block count, fixed summation order).
//   dense_axis, diag  : fast-diagonalisation solve of the separable part.
//   csr_axis          : Boehm prolongation / restriction along one axis.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <type_traits>

#include "tdz_internal.hpp"

namespace tdz {

namespace {

template <typename R>
__device__ __forceinline__ R pot_value(const PotParams &pp, R r2)
{
    if (pp.kind == 1) return R(0.5) * (R)pp.a0 * r2;
    if (pp.kind == 2) return -(R)pp.a0 / sqrt(r2 + (R)pp.a1);
    return R(0);
}

// 1/sqrt(r) for r > 0: the hardware estimate refined by one second-order
// Newton-Raphson (Householder) step, y = y0 + y0 e (1/2 + 3/8 e), e = 1 - r y0^2.
__device__ __forceinline__ double rsqrt_pos(double r)
{
    const double y0 = __builtin_amdgcn_rsq(r);
    const double e = fma(-r * y0, y0, 1.0);
    return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// single precision (preconditioner): the hardware estimate, about 1 ulp
__device__ __forceinline__ float rsqrt_pos(float r) { return __builtin_amdgcn_rsqf(r); }

// ------------------------------------------------------------------ potential
// One element per workgroup. LDS layout (doubles): tables Bx|By|Bz, then the
// complex stages c[n3][n2][n1], t1[n3][n2][Q1], t2[n3][Q2][Q1], F[Q3][Q2][Q1].
template <typename R>
__global__ void __launch_bounds__(128) pot_apply_generic_kernel(AxisDevT<R> a0, AxisDevT<R> a1, AxisDevT<R> a2, PotParams pp,
                                                                R gr, R gi, const cpx<R> *__restrict__ x,
                                                                cpx<R> *__restrict__ y)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    R *smem = reinterpret_cast<R *>(smem_raw);
    const int n1 = a0.p + 1, n2 = a1.p + 1, n3 = a2.p + 1;
    const int Q1 = a0.Q, Q2 = a1.Q, Q3 = a2.Q;
    const long eid = blockIdx.x;
    const int e1 = (int)(eid % a0.E);
    const int e2 = (int)((eid / a0.E) % a1.E);
    const int e3 = (int)(eid / ((long)a0.E * a1.E));
    const int o1 = a0.off[e1], o2 = a1.off[e2], o3 = a2.off[e3];
    const long g0 = a0.n, g1 = a1.n;

    R *Bx = smem, *By = Bx + n1 * Q1, *Bz = By + n2 * Q2;
    R *Xq = Bz + n3 * Q3, *Yq = Xq + Q1, *Zq = Yq + Q2;  // point coordinates
    R *Wx = Zq + Q3, *Wy = Wx + Q1, *Wz = Wy + Q2;        // weights
    size_t off = (size_t)(n1 * Q1 + n2 * Q2 + n3 * Q3 + 2 * (Q1 + Q2 + Q3));
    off = (off + 1) & ~(size_t)1;  // 16-byte alignment for the complex arrays
    cpx<R> *c = reinterpret_cast<cpx<R> *>(smem + off);
    cpx<R> *t1 = c + n1 * n2 * n3;
    cpx<R> *t2 = t1 + n3 * n2 * Q1;
    cpx<R> *F = t2 + n3 * Q2 * Q1;

    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < n1 * Q1; i += nt) Bx[i] = a0.B0[(size_t)e1 * n1 * Q1 + i];
    for (int i = tid; i < n2 * Q2; i += nt) By[i] = a1.B0[(size_t)e2 * n2 * Q2 + i];
    for (int i = tid; i < n3 * Q3; i += nt) Bz[i] = a2.B0[(size_t)e3 * n3 * Q3 + i];
    for (int i = tid; i < Q1; i += nt) { Xq[i] = a0.X[(size_t)e1 * Q1 + i]; Wx[i] = a0.W[(size_t)e1 * Q1 + i]; }
    for (int i = tid; i < Q2; i += nt) { Yq[i] = a1.X[(size_t)e2 * Q2 + i]; Wy[i] = a1.W[(size_t)e2 * Q2 + i]; }
    for (int i = tid; i < Q3; i += nt) { Zq[i] = a2.X[(size_t)e3 * Q3 + i]; Wz[i] = a2.W[(size_t)e3 * Q3 + i]; }
    for (int i = tid; i < n1 * n2 * n3; i += nt) {
        const int b1 = i % n1, b2 = (i / n1) % n2, b3 = i / (n1 * n2);
        c[i] = x[(o1 + b1) + g0 * ((o2 + b2) + g1 * (long)(o3 + b3))];
    }
    __syncthreads();
    // stage x: t1[(b3 n2 + b2) Q1 + q1]
    for (int i = tid; i < n3 * n2 * Q1; i += nt) {
        const int q1 = i % Q1, r = i / Q1;
        R sr = 0, si = 0;
        for (int b1 = 0; b1 < n1; ++b1) {
            const R bv = Bx[b1 * Q1 + q1];
            const cpx<R> v = c[r * n1 + b1];
            sr += bv * v.re; si += bv * v.im;
        }
        t1[i] = {sr, si};
    }
    __syncthreads();
    // stage y: t2[(b3 Q2 + q2) Q1 + q1]
    for (int i = tid; i < n3 * Q2 * Q1; i += nt) {
        const int q1 = i % Q1, q2 = (i / Q1) % Q2, b3 = i / (Q1 * Q2);
        R sr = 0, si = 0;
        for (int b2 = 0; b2 < n2; ++b2) {
            const R bv = By[b2 * Q2 + q2];
            const cpx<R> v = t1[(b3 * n2 + b2) * Q1 + q1];
            sr += bv * v.re; si += bv * v.im;
        }
        t2[i] = {sr, si};
    }
    __syncthreads();
    // stage z and the point coefficient g s V(r) w
    for (int i = tid; i < Q3 * Q2 * Q1; i += nt) {
        const int q1 = i % Q1, q2 = (i / Q1) % Q2, q3 = i / (Q1 * Q2);
        R sr = 0, si = 0;
        for (int b3 = 0; b3 < n3; ++b3) {
            const R bv = Bz[b3 * Q3 + q3];
            const cpx<R> v = t2[(b3 * Q2 + q2) * Q1 + q1];
            sr += bv * v.re; si += bv * v.im;
        }
        const R r2 = Xq[q1] * Xq[q1] + Yq[q2] * Yq[q2] + Zq[q3] * Zq[q3];
        const R cv = (R)pp.s * pot_value(pp, r2) * (Wx[q1] * Wy[q2] * Wz[q3]);
        const R cr = gr * cv, ci = gi * cv;
        F[i] = {cr * sr - ci * si, cr * si + ci * sr};
    }
    __syncthreads();
    // transposed z: t2[(b3 Q2 + q2) Q1 + q1] = sum_q3 Bz[b3,q3] F[(q3 Q2 + q2) Q1 + q1]
    for (int i = tid; i < n3 * Q2 * Q1; i += nt) {
        const int q21 = i % (Q2 * Q1), b3 = i / (Q2 * Q1);
        R sr = 0, si = 0;
        for (int q3 = 0; q3 < Q3; ++q3) {
            const R bv = Bz[b3 * Q3 + q3];
            const cpx<R> v = F[q3 * Q2 * Q1 + q21];
            sr += bv * v.re; si += bv * v.im;
        }
        t2[i] = {sr, si};
    }
    __syncthreads();
    // transposed y: t1[(b3 n2 + b2) Q1 + q1] = sum_q2 By[b2,q2] t2[(b3 Q2 + q2) Q1 + q1]
    for (int i = tid; i < n3 * n2 * Q1; i += nt) {
        const int q1 = i % Q1, b2 = (i / Q1) % n2, b3 = i / (Q1 * n2);
        R sr = 0, si = 0;
        for (int q2 = 0; q2 < Q2; ++q2) {
            const R bv = By[b2 * Q2 + q2];
            const cpx<R> v = t2[(b3 * Q2 + q2) * Q1 + q1];
            sr += bv * v.re; si += bv * v.im;
        }
        t1[i] = {sr, si};
    }
    __syncthreads();
    // transposed x and scatter-add
    for (int i = tid; i < n1 * n2 * n3; i += nt) {
        const int b1 = i % n1, r = i / n1;
        const int b2 = r % n2, b3 = r / n2;
        R sr = 0, si = 0;
        for (int q1 = 0; q1 < Q1; ++q1) {
            const R bv = Bx[b1 * Q1 + q1];
            const cpx<R> v = t1[r * Q1 + q1];
            sr += bv * v.re; si += bv * v.im;
        }
        R *dst = reinterpret_cast<R *>(y + ((o1 + b1) + g0 * ((o2 + b2) + g1 * (long)(o3 + b3))));
        unsafeAtomicAdd(dst, sr);
        unsafeAtomicAdd(dst + 1, si);
    }
}

// ------------------------------------------------------ potential, 2D tiles
// One workgroup owns TXE x TYE elements of a 2D grid with simple knots (element e
// starts at function e on both axes) and one thread per quadrature point:
//   load the (TXE+P) x (TYE+P) coefficient patch, x stage -> T1 [rows][x points],
//   y stage at every point times g s V(r) w -> F, transposed y stage -> T2, transposed
//   x stage -> patch contributions, atomically added to y (neighbouring tiles share P
//   functions per axis). Tile shapes are picked per grid by the autotuner.
template <typename R, int P, int Q, int TXE, int TYE, int KIND>
__global__ void __launch_bounds__(TXE * Q * TYE * Q) pot2d_kernel(AxisDevT<R> a0, AxisDevT<R> a1, PotParams pp,
                                                                  R gr, R gi, const cpx<R> *__restrict__ x,
                                                                  cpx<R> *__restrict__ y)
{
    constexpr int NB = P + 1, CX = TXE + P, CY = TYE + P, PX = TXE * Q, PY = TYE * Q, NT = PX * PY;
    __shared__ R Bx[TXE][NB][Q], By[TYE][NB][Q];
    __shared__ R Xq[PX], Wx[PX], Yq[PY], Wy[PY];
    __shared__ cpx<R> c[CY][CX];
    __shared__ cpx<R> T1[CY][PX];
    __shared__ cpx<R> F[PY][PX];
    const int tid = threadIdx.x;
    const int ex0 = blockIdx.x * TXE, ey0 = blockIdx.y * TYE;
    const int Ex = a0.E, Ey = a1.E;
    const long n0 = a0.n, n1 = a1.n;
    for (int i = tid; i < TXE * NB * Q; i += NT) {
        const int lx = i / (NB * Q), r = i % (NB * Q);
        Bx[lx][r / Q][r % Q] = ex0 + lx < Ex ? a0.B0[(size_t)(ex0 + lx) * NB * Q + r] : (R)0;
    }
    for (int i = tid; i < TYE * NB * Q; i += NT) {
        const int ly = i / (NB * Q), r = i % (NB * Q);
        By[ly][r / Q][r % Q] = ey0 + ly < Ey ? a1.B0[(size_t)(ey0 + ly) * NB * Q + r] : (R)0;
    }
    for (int i = tid; i < PX; i += NT) {
        const int e = ex0 + i / Q;
        Xq[i] = e < Ex ? a0.X[(size_t)e * Q + i % Q] : (R)0;
        Wx[i] = e < Ex ? a0.W[(size_t)e * Q + i % Q] : (R)0;   // zero weight: no contribution
    }
    for (int i = tid; i < PY; i += NT) {
        const int e = ey0 + i / Q;
        Yq[i] = e < Ey ? a1.X[(size_t)e * Q + i % Q] : (R)0;
        Wy[i] = e < Ey ? a1.W[(size_t)e * Q + i % Q] : (R)0;
    }
    for (int i = tid; i < CY * CX; i += NT) {
        const int jy = i / CX, ix = i % CX;
        const long gx = ex0 + ix, gy = ey0 + jy;
        c[jy][ix] = (gx < n0 && gy < n1) ? x[gx + n0 * gy] : cpx<R>{0, 0};
    }
    __syncthreads();
    // x stage: T1[j][px] = sum_b Bx[lx][b][q] c[j][lx + b]
    for (int i = tid; i < CY * PX; i += NT) {
        const int j = i / PX, px = i % PX, lx = px / Q, q = px % Q;
        R sr = 0, si = 0;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const R bv = Bx[lx][b][q];
            const cpx<R> v = c[j][lx + b];
            sr += bv * v.re;
            si += bv * v.im;
        }
        T1[j][px] = {sr, si};
    }
    __syncthreads();
    // y stage and the point coefficient g s V(r) w: one point per thread
    {
        const int py = tid / PX, px = tid % PX, ly = py / Q, q = py % Q;
        R sr = 0, si = 0;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const R bv = By[ly][b][q];
            const cpx<R> v = T1[ly + b][px];
            sr += bv * v.re;
            si += bv * v.im;
        }
        const R r2 = Xq[px] * Xq[px] + Yq[py] * Yq[py];
        R vv;
        if (KIND == 2) vv = -(R)pp.a0 / sqrt(r2 + (R)pp.a1);
        else vv = (R)0.5 * (R)pp.a0 * r2;
        const R cv = (R)pp.s * vv * Wx[px] * Wy[py];
        const R cr = gr * cv, ci = gi * cv;
        F[py][px] = {cr * sr - ci * si, cr * si + ci * sr};
    }
    __syncthreads();
    // transposed y: T1[j][px] = sum_{ly + b = j} sum_q By[ly][b][q] F[ly Q + q][px]
    for (int i = tid; i < CY * PX; i += NT) {
        const int j = i / PX, px = i % PX;
        R sr = 0, si = 0;
#pragma unroll
        for (int ly = 0; ly < TYE; ++ly) {
            const int b = j - ly;
            if (b < 0 || b > P) continue;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const R bv = By[ly][b][q];
                const cpx<R> v = F[ly * Q + q][px];
                sr += bv * v.re;
                si += bv * v.im;
            }
        }
        T1[j][px] = {sr, si};
    }
    __syncthreads();
    // transposed x and scatter-add: y[patch (j, i)] += sum_{lx + b = i} sum_q Bx[lx][b][q] T1[j][lx Q + q]
    for (int k = tid; k < CY * CX; k += NT) {
        const int j = k / CX, i = k % CX;
        const long gx = ex0 + i, gy = ey0 + j;
        if (gx >= n0 || gy >= n1) continue;
        R sr = 0, si = 0;
#pragma unroll
        for (int lx = 0; lx < TXE; ++lx) {
            const int b = i - lx;
            if (b < 0 || b > P) continue;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const R bv = Bx[lx][b][q];
                const cpx<R> v = T1[j][lx * Q + q];
                sr += bv * v.re;
                si += bv * v.im;
            }
        }
        R *dst = reinterpret_cast<R *>(y + (gx + n0 * gy));
        unsafeAtomicAdd(dst, sr);
        unsafeAtomicAdd(dst + 1, si);
    }
}

// ------------------------------------------------------ potential, 3D marching
// One workgroup owns a TE x TE tile of elements in (x, y) and marches through
// all element layers in z. The (x, y) stages of the sum factorisation are done
// once per coefficient plane (not once per element), each thread keeps the z
// lines of CPT fixed (qx, qy) point columns in registers: a rolling window of
// the P+1 input planes an element needs and of the P+1 output planes it
// touches. Per element layer:
//   A  load plane ez+P, x stage -> T1 (LDS), y stage -> register window
//   B  z stage at the Q points of element ez, times s V(r) w, transposed z
//      stage into the output window
//   C  output plane ez is complete: transposed y then x stages through LDS,
//      scaled by g, stored to a scratch plane of the tile's offset class
//      (neighbouring tiles share p planes of functions in x and y; pot3d_sum_kernel
//      adds the classes into y in a fixed order, so the result is deterministic).
// Requires simple interior knots (element e starts at function e) and dim 3.
template <typename R, int P, int Q, int TE, int CPT, int KIND>
__global__ void __launch_bounds__((TE * Q) * (TE * Q) / CPT)
pot3d_kernel(AxisDevT<R> ax, AxisDevT<R> ay, AxisDevT<R> az, PotParams pp, R gr, R gi,
             const cpx<R> *__restrict__ x, cpx<R> *__restrict__ scr)
{
    constexpr int NP = P + 1, TQ = TE * Q, NCOL = TQ * TQ, NT = NCOL / CPT, TW = TE + P;
    __shared__ cpx<R> cin[TW * TW];   // input plane  [jy][jx]
    __shared__ cpx<R> T1[TW * TQ];    // x stage      [jy][qx]
    __shared__ cpx<R> G2p[NCOL];      // output plane at the points [qy][qx]
    __shared__ cpx<R> H1[TW * TQ];    // transposed y [jy][qx]
    __shared__ R Bxs[TE * NP * Q], Bys[TE * NP * Q];

    const int tid = threadIdx.x;
    const int ex0 = blockIdx.x * TE, ey0 = blockIdx.y * TE;
    const int n0 = ax.n, n1 = ay.n, Ez = az.E;
    const long plane = (long)n0 * n1;

    for (int i = tid; i < TE * NP * Q; i += NT) {
        const int el = i / (NP * Q), r = i % (NP * Q);
        Bxs[i] = ex0 + el < ax.E ? ax.B0[(long)(ex0 + el) * NP * Q + r] : 0.0;
        Bys[i] = ey0 + el < ay.E ? ay.B0[(long)(ey0 + el) * NP * Q + r] : 0.0;
    }
    // per column constants: soft-core s V w = cst / sqrt(xy2 + z^2) wz with xy2 = x^2 + y^2 + a;
    // harmonic s V w = cst (xy2 + z^2) wz with xy2 = x^2 + y^2
    R xy2[CPT], wxy[CPT], By[CPT][NP];
    bool valid[CPT];
    int cqx[CPT], cey[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int col = tid + c * NT;
        const int qx = col % TQ, qy = col / TQ;
        const int exl = qx / Q, q1 = qx % Q, eyl = qy / Q, q2 = qy % Q;
        const int ex = ex0 + exl, ey = ey0 + eyl;
        valid[c] = ex < ax.E && ey < ay.E;
        cqx[c] = qx;
        cey[c] = eyl;
        const int exs = valid[c] ? ex : 0, eys = valid[c] ? ey : 0;
        const R xv = ax.X[(long)exs * Q + q1], yv = ay.X[(long)eys * Q + q2];
        const R w = ax.W[(long)exs * Q + q1] * ay.W[(long)eys * Q + q2];
        xy2[c] = xv * xv + yv * yv + (KIND == 2 ? pp.a1 : 0.0);
        wxy[c] = KIND == 2 ? -pp.s * pp.a0 * w : 0.5 * pp.s * pp.a0 * w;
#pragma unroll
        for (int a = 0; a < NP; ++a) By[c][a] = ay.B0[((long)eys * NP + a) * Q + q2];
    }
    R tw[CPT][NP][2];   // input window: planes ez .. ez+P at my columns
    R gw[CPT][NP][2];   // output window
#pragma unroll
    for (int c = 0; c < CPT; ++c)
#pragma unroll
        for (int a = 0; a < NP; ++a) tw[c][a][0] = tw[c][a][1] = gw[c][a][0] = gw[c][a][1] = 0.0;

    // A: plane jz -> window slot `slot`
    auto stage_in = [&](int jz, int slot) {
        for (int i = tid; i < TW * TW; i += NT) {
            const int jx = ex0 + i % TW, jy = ey0 + i / TW;
            cin[i] = (jx < n0 && jy < n1) ? x[jx + (long)n0 * jy + plane * jz] : cpx<R>{0.0, 0.0};
        }
        __syncthreads();
        for (int i = tid; i < TW * TQ; i += NT) {
            const int qx = i % TQ, jy = i / TQ, exl = qx / Q, q1 = qx % Q;
            R sr = 0.0, si = 0.0;
#pragma unroll
            for (int a = 0; a < NP; ++a) {
                const R b = Bxs[(exl * NP + a) * Q + q1];
                const cpx<R> v = cin[jy * TW + exl + a];
                sr += b * v.re;
                si += b * v.im;
            }
            T1[i] = {sr, si};
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            R sr = 0.0, si = 0.0;
#pragma unroll
            for (int a = 0; a < NP; ++a) {
                const cpx<R> v = T1[(cey[c] + a) * TQ + cqx[c]];
                sr += By[c][a] * v.re;
                si += By[c][a] * v.im;
            }
            tw[c][slot][0] = sr;
            tw[c][slot][1] = si;
        }
    };

    // C: emit output plane jz from ring slot `slot` (then cleared: it becomes plane jz + P + 1)
    auto emit = [&](int jz, int slot) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const R r = gw[c][slot][0], i = gw[c][slot][1];
            G2p[tid + c * NT] = {gr * r - gi * i, gr * i + gi * r};
            gw[c][slot][0] = gw[c][slot][1] = 0.0;
        }
        __syncthreads();
        for (int i = tid; i < TW * TQ; i += NT) {
            const int qx = i % TQ, jy = i / TQ;
            R sr = 0.0, si = 0.0;
#pragma unroll 1
            for (int eyl = 0; eyl < TE; ++eyl) {
                const int a = jy - eyl;
                if (a < 0 || a > P) continue;
#pragma unroll
                for (int q2 = 0; q2 < Q; ++q2) {
                    const R b = Bys[(eyl * NP + a) * Q + q2];
                    const cpx<R> v = G2p[(eyl * Q + q2) * TQ + qx];
                    sr += b * v.re;
                    si += b * v.im;
                }
            }
            H1[i] = {sr, si};
        }
        __syncthreads();
        for (int i = tid; i < TW * TW; i += NT) {
            const int jxl = i % TW, jyl = i / TW;
            const int gx = ex0 + jxl, gy = ey0 + jyl;
            if (gx >= n0 || gy >= n1) continue;
            R sr = 0.0, si = 0.0;
#pragma unroll 1
            for (int exl = 0; exl < TE; ++exl) {
                const int a = jxl - exl;
                if (a < 0 || a > P) continue;
#pragma unroll
                for (int q1 = 0; q1 < Q; ++q1) {
                    const R b = Bxs[(exl * NP + a) * Q + q1];
                    const cpx<R> v = H1[jyl * TQ + exl * Q + q1];
                    sr += b * v.re;
                    si += b * v.im;
                }
            }
            // the tile's contribution to function (gx, gy, jz) goes to the scratch plane of its
            // offset class (jxl / TE, jyl / TE): every (function, class) has one writer
            constexpr int MW = (TW + TE - 1) / TE;
            scr[((long)((jyl / TE) * MW + jxl / TE) * az.n + jz) * plane + gx + (long)n0 * gy] = {sr, si};
        }
        // no barrier here: the next writes of H1 and G2p follow the next emit's first
        // barrier, those of cin and the z tables (stage_in) only replace data whose
        // last readers ran before this emit's barriers
    };

    // the z windows are rings of NP planes (plane j in slot j mod NP): with the layer loop
    // unrolled by NP every slot index is a compile-time constant and no window is shifted
#pragma unroll
    for (int s = 0; s < P; ++s) stage_in(s, s);
    for (int e0 = 0; e0 < Ez + P; e0 += NP) {
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int ez = e0 + u;   // block-uniform
            if (ez >= Ez + P) break;
            if (ez < Ez) {
                stage_in(ez + P, (u + P) % NP);
                // B: z stage, coefficient, transposed z stage. The z tables of element ez are the same
                // for every lane: read from global memory at a wave-uniform address they become scalar
                // loads and SGPR operands of the FMAs (no LDS traffic for them)
                typedef const __attribute__((address_space(4))) R *cptr;   // constant address space: invariant
                const cptr bz = (cptr)(az.B0 + (long)ez * NP * Q);
                const cptr zq = (cptr)(az.X + (long)ez * Q);
                const cptr wq = (cptr)(az.W + (long)ez * Q);
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    if (!valid[c]) continue;
#pragma unroll 2
                    for (int qz = 0; qz < Q; ++qz) {
                        R fr = 0.0, fi = 0.0;
#pragma unroll
                        for (int a = 0; a < NP; ++a) {
                            const R b = bz[a * Q + qz];
                            fr += b * tw[c][(u + a) % NP][0];
                            fi += b * tw[c][(u + a) % NP][1];
                        }
                        const R r2 = xy2[c] + zq[qz] * zq[qz];
                        const R v = (KIND == 2 ? rsqrt_pos(r2) : r2) * (wxy[c] * wq[qz]);
                        fr *= v;
                        fi *= v;
#pragma unroll
                        for (int a = 0; a < NP; ++a) {
                            const R b = bz[a * Q + qz];
                            gw[c][(u + a) % NP][0] += b * fr;
                            gw[c][(u + a) % NP][1] += b * fi;
                        }
                    }
                }
            }
            emit(ez, u);
        }
    }
}

// y[I] += the contributions of the up to MW^2 tiles that cover function I, summed in a
// fixed order (class y outer, class x inner). Tile bx covers functions bx TE .. bx TE + TW - 1.
template <typename R>
__global__ void pot3d_sum_kernel(int n0, int n1, int n2, int TE, int TW, int nbx, int nby,
                                 const cpx<R> *__restrict__ scr, cpx<R> *__restrict__ y)
{
    const long plane = (long)n0 * n1, N = plane * n2;
    const int MW = (TW + TE - 1) / TE;
    for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < N; I += (long)gridDim.x * blockDim.x) {
        const int gx = (int)(I % n0), gy = (int)((I / n0) % n1);
        R sr = 0, si = 0;
        for (int my = 0; my < MW; ++my) {
            const int jyl = my * TE + gy % TE, by = gy / TE - my;
            if (by < 0 || by >= nby || jyl >= TW) continue;
            for (int mx = 0; mx < MW; ++mx) {
                const int jxl = mx * TE + gx % TE, bx = gx / TE - mx;
                if (bx < 0 || bx >= nbx || jxl >= TW) continue;
                const cpx<R> v = scr[(long)(my * MW + mx) * N + I];
                sr += v.re;
                si += v.im;
            }
        }
        cpx<R> o = y[I];
        o.re += sr;
        o.im += si;
        y[I] = o;
    }
}

// ------------------------------------------------------------------ Jacobi-Chebyshev smoother
// dinv[I] = 1 / diag(A)[I], A = the level operator with its current shift: the Kronecker
// terms' diagonal from the band diagonals (B_d complex, M_d real), plus g times the
// potential's diagonal (Vc: the Galerkin stencil's centre entries, offset-major with
// half-widths p; or Vl: s V at the Greville points times diag(M), the finest level's
// lumped value; or none). Wall rows are the identity.
template <typename R>
__global__ void jacobi_diag_kernel(int n0, int n1, int n2, int p0, int p1, int p2, const R *__restrict__ M0,
                                   const R *__restrict__ M1, const R *__restrict__ M2, const cpx<R> *__restrict__ B0,
                                   const cpx<R> *__restrict__ B1, const cpx<R> *__restrict__ B2,
                                   const R *__restrict__ Vc, const R *__restrict__ Vl, R gr, R gi,
                                   const unsigned char *__restrict__ w0, const unsigned char *__restrict__ w1,
                                   const unsigned char *__restrict__ w2, cpx<R> *__restrict__ dinv)
{
    const long N = (long)n0 * n1 * n2;
    for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < N; I += (long)gridDim.x * blockDim.x) {
        const int i = (int)(I % n0), j = (int)((I / n0) % n1), k = (int)(I / ((long)n0 * n1));
        if (w0[i] || w1[j] || w2[k]) {
            dinv[I] = {1, 0};
            continue;
        }
        const R mx = M0[(long)i * (2 * p0 + 1) + p0], my = M1[(long)j * (2 * p1 + 1) + p1],
                mz = M2[(long)k * (2 * p2 + 1) + p2];
        const cpx<R> bx = B0[(long)i * (2 * p0 + 1) + p0], by = B1[(long)j * (2 * p1 + 1) + p1],
                     bz = B2[(long)k * (2 * p2 + 1) + p2];
        R dr = bx.re * my * mz + mx * by.re * mz + mx * my * bz.re;
        R di = bx.im * my * mz + mx * by.im * mz + mx * my * bz.im;
        R v = 0;
        if (Vc) {   // centre entry (offsets p0, p1, p2) of the offset-major stencil
            const long D0 = 2 * p0 + 1, D1 = 2 * p1 + 1;
            v = Vc[((((long)p2 * n2 + k) * D1 + p1) * n1 + j) * D0 * n0 + (long)p0 * n0 + i];
        } else if (Vl) {
            v = Vl[I] * mx * my * mz;
        }
        dr += gr * v;
        di += gi * v;
        const R den = dr * dr + di * di;
        dinv[I] = {dr / den, -di / den};
    }
}

// d = a d + b (dinv .* res) (a = 0: d is written without being read); e += d when e
template <typename R>
__global__ void cheb_update_kernel(long N, R a, R b, const cpx<R> *__restrict__ dinv, const cpx<R> *__restrict__ res,
                                   cpx<R> *__restrict__ d, cpx<R> *__restrict__ e)
{
    for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < N; I += (long)gridDim.x * blockDim.x) {
        const cpx<R> q = dinv[I], r = res[I];
        cpx<R> v = {b * (q.re * r.re - q.im * r.im), b * (q.re * r.im + q.im * r.re)};
        if (a != R(0)) {
            const cpx<R> o = d[I];
            v.re += a * o.re;
            v.im += a * o.im;
        }
        d[I] = v;
        if (e) {
            cpx<R> t = e[I];
            t.re += v.re;
            t.im += v.im;
            e[I] = t;
        }
    }
}

// ------------------------------------------------------------------ separable
// x pass: t = Mx x, u = Bx x (one thread per output)
template <typename R>
__global__ void sep_x_kernel(int n0, long lines, int p, const R *__restrict__ M, const cpx<R> *__restrict__ B,
                             const cpx<R> *__restrict__ x, cpx<R> *__restrict__ t, cpx<R> *__restrict__ u)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= lines * n0) return;
    const int i = (int)(I % n0);
    const long base = I - i;
    const int bw = 2 * p + 1;
    const int lo = i - p < 0 ? 0 : i - p, hi = i + p > n0 - 1 ? n0 - 1 : i + p;
    R tr = 0, ti = 0, ur = 0, ui = 0;
    for (int J = lo; J <= hi; ++J) {
        const cpx<R> v = x[base + J];
        const R m = M[i * bw + (J - i + p)];
        const cpx<R> b = B[i * bw + (J - i + p)];
        tr += m * v.re; ti += m * v.im;
        ur += b.re * v.re - b.im * v.im;
        ui += b.re * v.im + b.im * v.re;
    }
    t[I] = {tr, ti};
    u[I] = {ur, ui};
}

// y pass: a = My u + By t, b = My t
template <typename R>
__global__ void sep_y_kernel(int n0, int n1, long N, int p, const R *__restrict__ M,
                             const cpx<R> *__restrict__ B, const cpx<R> *__restrict__ t,
                             const cpx<R> *__restrict__ u, cpx<R> *__restrict__ a, cpx<R> *__restrict__ b)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int i = (int)(I % n0);
    const int j = (int)((I / n0) % n1);
    const long k = I / ((long)n0 * n1);
    const int bw = 2 * p + 1;
    const int lo = j - p < 0 ? 0 : j - p, hi = j + p > n1 - 1 ? n1 - 1 : j + p;
    R ar = 0, ai = 0, br = 0, bi = 0;
    for (int J = lo; J <= hi; ++J) {
        const long src = i + (long)n0 * (J + (long)n1 * k);
        const cpx<R> tv = t[src], uv = u[src];
        const R m = M[j * bw + (J - j + p)];
        const cpx<R> bb = B[j * bw + (J - j + p)];
        ar += m * uv.re + bb.re * tv.re - bb.im * tv.im;
        ai += m * uv.im + bb.re * tv.im + bb.im * tv.re;
        br += m * tv.re; bi += m * tv.im;
    }
    a[I] = {ar, ai};
    b[I] = {br, bi};
}

// z pass: y = Mz a + Bz b
template <typename R>
__global__ void sep_z_kernel(long plane, int n2, int p, const R *__restrict__ M, const cpx<R> *__restrict__ B,
                             const cpx<R> *__restrict__ a, const cpx<R> *__restrict__ b, cpx<R> *__restrict__ y)
{
    // M: the z factor of the x and y terms (the mass band, or the 1/m-weighted one)
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= plane * n2) return;
    const long m = I % plane;
    const int k = (int)(I / plane);
    const int bw = 2 * p + 1;
    const int lo = k - p < 0 ? 0 : k - p, hi = k + p > n2 - 1 ? n2 - 1 : k + p;
    R yr = 0, yi = 0;
    for (int K = lo; K <= hi; ++K) {
        const cpx<R> av = a[m + plane * K], bv = b[m + plane * K];
        const R mm = M[k * bw + (K - k + p)];
        const cpx<R> bb = B[k * bw + (K - k + p)];
        yr += mm * av.re + bb.re * bv.re - bb.im * bv.im;
        yi += mm * av.im + bb.re * bv.im + bb.im * bv.re;
    }
    y[I] = {yr, yi};
}

// Fused separable apply (FP64, 3D, equal orders P on the three axes), z-marching:
//   y = Mz' My Bx x + Mz' By Mx x + Bz My Mx x
// (Mz' the z factor of the x and y terms). A block owns a 16 x 16 tile of (i, j)
// columns and a chunk of output planes k; it walks the input planes K that reach the
// chunk. Per plane: the x band pass for the tile's columns on the 16 + 2P rows it
// needs (plane tile with halo staged in LDS), then the y band pass for each (i, j), and
// the z band scatters a = My u + By t, b = My t into a ring of 2P + 1 output planes
// held in registers (slot of output k: (k + P) mod (2P + 1), compile-time in the
// unrolled ring loop); output k - P is complete after plane K and is stored. One read of
// x and one write of y per point, instead of the three passes' six vector transfers.
template <int P>
__global__ void __launch_bounds__(256) sep3_kernel(int n0, int n1, int n2, int kchunk, const double *__restrict__ Mx,
                                                   const dcomplex *__restrict__ Bx, const double *__restrict__ My,
                                                   const dcomplex *__restrict__ By, const double *__restrict__ Mz,
                                                   const dcomplex *__restrict__ Bz, const dcomplex *__restrict__ x,
                                                   dcomplex *__restrict__ y)
{
    constexpr int R = 2 * P + 1, T = 16, H = T + 2 * P;
    __shared__ dcomplex Xs[H * H];            // plane tile with halo, [jj][ii]
    __shared__ dcomplex Ts[H * T], Us[H * T];  // x pass on the halo rows, [jj][i]
    const int tid = threadIdx.x, ti = tid % T, tj = tid / T;
    const int tilesx = (n0 + T - 1) / T, tilesy = (n1 + T - 1) / T;
    int b = blockIdx.x;
    const int bx = b % tilesx;
    b /= tilesx;
    const int by = b % tilesy, bz = b / tilesy;
    const int i0 = bx * T, j0 = by * T, k0 = bz * kchunk, k1 = min(k0 + kchunk, n2);
    const int i = i0 + ti, j = j0 + tj;
    const bool own = i < n0 && j < n1;
    const int ic = min(i, n0 - 1), jc = min(j, n1 - 1);
    double mx[R], my[R];
    dcomplex bxv[R], byv[R], acc[R];
#pragma unroll
    for (int l = 0; l < R; ++l) {
        mx[l] = Mx[ic * R + l];
        bxv[l] = Bx[ic * R + l];
        my[l] = My[jc * R + l];
        byv[l] = By[jc * R + l];
        acc[l] = {0.0, 0.0};
    }
    const long plane = (long)n0 * n1;
    const int Kbeg = max(k0 - P, 0), Kend = min(k1 + P, n2);
    // the plane tile of input plane K, each thread's share in registers: the next
    // plane's loads are in flight while the current plane is computed
    constexpr int NX = (H * H + 255) / 256;
    dcomplex xv[NX];
    auto load_plane = [&](int K) {
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int e = tid + 256 * q;
            const int ii = i0 - P + e % H, jj = j0 - P + e / H;
            xv[q] = {0.0, 0.0};
            if (e < H * H && ii >= 0 && ii < n0 && jj >= 0 && jj < n1) xv[q] = x[K * plane + (long)jj * n0 + ii];
        }
    };
    load_plane(Kbeg);
    for (int Kb = (Kbeg / R) * R; Kb < k1 + P; Kb += R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int K = Kb + r;
            if (K >= Kbeg && K < Kend) {   // block-uniform
#pragma unroll
                for (int q = 0; q < NX; ++q) {
                    const int e = tid + 256 * q;
                    if (e < H * H) Xs[e] = xv[q];
                }
                __syncthreads();
                if (K + 1 < Kend) load_plane(K + 1);
                for (int e = tid; e < H * T; e += 256) {   // e % T == ti: the same x band row
                    const int jj = e / T;
                    double tr = 0, tim = 0, ur = 0, ui = 0;
#pragma unroll
                    for (int l = 0; l < R; ++l) {
                        const dcomplex v = Xs[jj * H + ti + l];
                        tr += mx[l] * v.re;
                        tim += mx[l] * v.im;
                        ur += bxv[l].re * v.re - bxv[l].im * v.im;
                        ui += bxv[l].re * v.im + bxv[l].im * v.re;
                    }
                    Ts[e] = {tr, tim};
                    Us[e] = {ur, ui};
                }
                __syncthreads();
                double ar = 0, ai = 0, br = 0, bi = 0;
#pragma unroll
                for (int l = 0; l < R; ++l) {
                    const dcomplex tv = Ts[(tj + l) * T + ti], uv = Us[(tj + l) * T + ti];
                    ar += my[l] * uv.re + byv[l].re * tv.re - byv[l].im * tv.im;
                    ai += my[l] * uv.im + byv[l].re * tv.im + byv[l].im * tv.re;
                    br += my[l] * tv.re;
                    bi += my[l] * tv.im;
                }
#pragma unroll
                for (int d = -P; d <= P; ++d) {   // output k = K + d, band entry (k, K) = P - d
                    const int k = K + d;
                    if (k >= k0 && k < k1) {
                        const double mz = Mz[k * R + P - d];
                        const dcomplex bzv = Bz[k * R + P - d];
                        dcomplex &o = acc[(r + d + P) % R];
                        o.re += mz * ar + bzv.re * br - bzv.im * bi;
                        o.im += mz * ai + bzv.re * bi + bzv.im * br;
                    }
                }
            }
            const int ko = K - P;   // complete: every input plane within P of it is done
            if (ko >= k0 && ko < k1 && own) y[ko * plane + (long)j * n0 + i] = acc[r];
            acc[r] = {0.0, 0.0};
        }
    }
}

template <typename R>
__global__ void wall_fix_kernel(int n0, int n1, long N, const unsigned char *w0, const unsigned char *w1,
                                const unsigned char *w2, const cpx<R> *__restrict__ x, cpx<R> *__restrict__ y)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int i = (int)(I % n0), j = (int)((I / n0) % n1);
    const long k = I / ((long)n0 * n1);
    if (w0[i] | w1[j] | w2[k]) y[I] = x[I];
}

// ------------------------------------------------------------- reductions
constexpr int kRedThreads = 256;

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh)
{
    // wave64 butterfly then across waves; fixed order => deterministic
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    __syncthreads();
    return s;
}

// partial[(i * gridDim.x) + block] = sum over the block's chunk of conj(V_i) w, K vectors per pass
template <int K>
__global__ void __launch_bounds__(kRedThreads) multi_dot_kernel(const dcomplex *__restrict__ V, long ld, int i0,
                                                                int nv, const dcomplex *__restrict__ w, long N,
                                                                dcomplex *__restrict__ partial)
{
    __shared__ double sh[2 * K][kRedThreads / 64];
    double acc[2 * K];
#pragma unroll
    for (int k = 0; k < 2 * K; ++k) acc[k] = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const dcomplex wv = w[I];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (i0 + k < nv) {
                const dcomplex v = V[(long)(i0 + k) * ld + I];
                acc[2 * k] += v.re * wv.re + v.im * wv.im;
                acc[2 * k + 1] += v.re * wv.im - v.im * wv.re;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 2 * K; ++k) {
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        if (i0 + k < nv) {
            double sr = 0, si = 0;
            for (int w2 = 0; w2 < kRedThreads / 64; ++w2) { sr += sh[2 * k][w2]; si += sh[2 * k + 1][w2]; }
            partial[(long)(i0 + k) * gridDim.x + blockIdx.x] = {sr, si};
        }
    }
}

// out[i] = sum_b partial[i * nb + b], one block per i, fixed order
__global__ void reduce_complex_kernel(const dcomplex *__restrict__ partial, int nb, dcomplex *__restrict__ out)
{
    __shared__ double sh[kRedThreads / 64];
    const int i = blockIdx.x;
    double sr = 0, si = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        sr += partial[(long)i * nb + b].re;
        si += partial[(long)i * nb + b].im;
    }
    sr = block_sum(sr, sh);
    si = block_sum(si, sh);
    if (threadIdx.x == 0) out[i] = {sr, si};
}

// out[i] = sum_b partial[i nb + b], one block per i, fixed order
__global__ void reduce_real_kernel(const double *__restrict__ partial, int nb, double *__restrict__ out)
{
    __shared__ double sh[kRedThreads / 64];
    const double *p = partial + (long)blockIdx.x * nb;
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += p[b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedThreads) multi_axpy_norm_kernel(const dcomplex *__restrict__ V, long ld, int nv,
                                                                      const dcomplex *__restrict__ h,
                                                                      dcomplex *__restrict__ w, long N,
                                                                      double *__restrict__ partial)
{
    __shared__ dcomplex hs[64];
    __shared__ double sh[kRedThreads / 64];
    for (int i = threadIdx.x; i < nv; i += blockDim.x) hs[i] = h[i];
    __syncthreads();
    double acc = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        dcomplex v = w[I];
        for (int i = 0; i < nv; ++i) {
            const dcomplex b = V[(long)i * ld + I];
            v.re -= hs[i].re * b.re - hs[i].im * b.im;
            v.im -= hs[i].re * b.im + hs[i].im * b.re;
        }
        w[I] = v;
        acc += v.re * v.re + v.im * v.im;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// The first Gram-Schmidt pass fused with the projections of its result (the second pass's
// dot products): w -= sum_i h_i V_i (nv <= K), with the block sums of ||w||^2 (pnorm[b])
// and of conj(V_i) . w (pdot[i nb + b]); each V_i is read once.
template <int K>
__global__ void __launch_bounds__(kRedThreads) multi_axpy_norm_dot_kernel(const dcomplex *__restrict__ V, long ld,
                                                                          int nv, const dcomplex *__restrict__ h,
                                                                          dcomplex *__restrict__ w, long N,
                                                                          double *__restrict__ pnorm,
                                                                          dcomplex *__restrict__ pdot)
{
    __shared__ dcomplex hs[K];
    __shared__ double sh[2 * K + 1][kRedThreads / 64];
    if ((int)threadIdx.x < K) hs[threadIdx.x] = (int)threadIdx.x < nv ? h[threadIdx.x] : dcomplex{0.0, 0.0};
    __syncthreads();
    double acc = 0, dr[K], di[K];
#pragma unroll
    for (int k = 0; k < K; ++k) dr[k] = di[k] = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        dcomplex vb[K];
        dcomplex v = w[I];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < nv) {
                vb[k] = V[(long)k * ld + I];
                v.re -= hs[k].re * vb[k].re - hs[k].im * vb[k].im;
                v.im -= hs[k].re * vb[k].im + hs[k].im * vb[k].re;
            }
        }
        w[I] = v;
        acc += v.re * v.re + v.im * v.im;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < nv) {
                dr[k] += vb[k].re * v.re + vb[k].im * v.im;
                di[k] += vb[k].re * v.im - vb[k].im * v.re;
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) sh[2 * K][wave] = acc;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double a = dr[k], b = di[k];
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
        }
        if (lane == 0) { sh[2 * k][wave] = a; sh[2 * k + 1][wave] = b; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w2 = 0; w2 < kRedThreads / 64; ++w2) s += sh[2 * K][w2];
        pnorm[blockIdx.x] = s;
    }
    if ((int)threadIdx.x < nv) {
        const int k = threadIdx.x;
        double sr = 0, si = 0;
        for (int w2 = 0; w2 < kRedThreads / 64; ++w2) { sr += sh[2 * k][w2]; si += sh[2 * k + 1][w2]; }
        pdot[(long)k * gridDim.x + blockIdx.x] = {sr, si};
    }
}

// w = s (w - sum_i h_i V_i), h on device: the deferred second Gram-Schmidt correction and
// the normalisation in one pass
__global__ void correct_scale_kernel(const dcomplex *__restrict__ V, long ld, int nv, const dcomplex *__restrict__ h,
                                     dcomplex *__restrict__ w, long N, double s)
{
    __shared__ dcomplex hs[64];
    for (int i = threadIdx.x; i < nv; i += blockDim.x) hs[i] = h[i];
    __syncthreads();
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        dcomplex v = w[I];
        for (int i = 0; i < nv; ++i) {
            const dcomplex b = V[(long)i * ld + I];
            v.re -= hs[i].re * b.re - hs[i].im * b.im;
            v.im -= hs[i].re * b.im + hs[i].im * b.re;
        }
        w[I] = {s * v.re, s * v.im};
    }
}

__global__ void multi_axpy_kernel(const dcomplex *__restrict__ Z, long ld, int nv, const dcomplex *__restrict__ c,
                                  const dcomplex *base, dcomplex *x, long N, CoefArg cv)
{
    __shared__ dcomplex cs[kCoefArgMax];
    for (int i = threadIdx.x; i < nv; i += blockDim.x) cs[i] = c ? c[i] : cv.v[i];
    __syncthreads();
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        dcomplex v = base[I];
        for (int i = 0; i < nv; ++i) {
            const dcomplex b = Z[(long)i * ld + I];
            v.re += cs[i].re * b.re - cs[i].im * b.im;
            v.im += cs[i].re * b.im + cs[i].im * b.re;
        }
        x[I] = v;
    }
}

// y = a x + b y; with b = 0 y is not read (BLAS semantics: y may hold uninitialised memory)
__global__ void axpby_kernel(dcomplex a, const dcomplex *__restrict__ x, dcomplex b, dcomplex *__restrict__ y, long N)
{
    const long stride = (long)gridDim.x * blockDim.x;
    const bool ry = b.re != 0.0 || b.im != 0.0;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const dcomplex xv = x[I], yv = ry ? y[I] : dcomplex{0.0, 0.0};
        y[I] = {a.re * xv.re - a.im * xv.im + b.re * yv.re - b.im * yv.im,
                a.re * xv.im + a.im * xv.re + b.re * yv.im + b.im * yv.re};
    }
}

// y = x / sqrt(n2) (in place allowed; a zero norm leaves zero)
__global__ void scale_invsqrt_kernel(const dcomplex *x, const double *__restrict__ n2, dcomplex *y, long N)
{
    const double s = n2[0] > 0.0 ? 1.0 / sqrt(n2[0]) : 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const dcomplex v = x[I];
        y[I] = {v.re * s, v.im * s};
    }
}

__global__ void __launch_bounds__(kRedThreads) norm2_kernel(const dcomplex *__restrict__ x, long N,
                                                            double *__restrict__ partial)
{
    __shared__ double sh[kRedThreads / 64];
    double acc = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const dcomplex v = x[I];
        acc += v.re * v.re + v.im * v.im;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

template <typename R>
__global__ void axpby_t_kernel(cpx<R> a, const cpx<R> *__restrict__ x, cpx<R> b, cpx<R> *__restrict__ y, long N)
{
    const long stride = (long)gridDim.x * blockDim.x;
    const bool ry = b.re != R(0) || b.im != R(0);   // b = 0: y is not read
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const cpx<R> xv = x[I], yv = ry ? y[I] : cpx<R>{R(0), R(0)};
        y[I] = {a.re * xv.re - a.im * xv.im + b.re * yv.re - b.im * yv.im,
                a.re * xv.im + a.im * xv.re + b.re * yv.im + b.im * yv.re};
    }
}

// y (complex64) = x (complex128), rounded
__global__ void demote_kernel(const dcomplex *__restrict__ x, fcomplex *__restrict__ y, long N)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const dcomplex v = x[I];
        y[I] = {(float)v.re, (float)v.im};
    }
}

// y (complex128) = x (complex64)
__global__ void promote_kernel(const fcomplex *__restrict__ x, dcomplex *__restrict__ y, long N)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const fcomplex v = x[I];
        y[I] = {(double)v.re, (double)v.im};
    }
}

// q = r - g Vg .* My - c - n (n optional): the right-hand side of the second
// separable solve of the first-order corrected preconditioner
__global__ void corr_combine_kernel(const fcomplex *__restrict__ r, const fcomplex *__restrict__ My,
                                    const fcomplex *c, const fcomplex *__restrict__ n,
                                    const float *__restrict__ Vg, fcomplex g, fcomplex *q, long N,
                                    const float *__restrict__ no, fcomplex ng)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const fcomplex m = My[I], cv = c[I], rv = r ? r[I] : fcomplex{0.f, 0.f};
        const float v = Vg[I];
        float qr = rv.re - v * (g.re * m.re - g.im * m.im) - cv.re;
        float qi = rv.im - v * (g.re * m.im + g.im * m.re) - cv.im;
        if (n) { qr -= n[I].re; qi -= n[I].im; }
        if (no) {
            const float ore = no[I], oim = no[N + I];
            qr -= ng.re * ore - ng.im * oim;
            qi -= ng.re * oim + ng.im * ore;
        }
        q[I] = {qr, qi};
    }
}

// --------------------------------------------------------------- dense axis
// y[.., i, ..] = sum_k T(i, k) x[.., k, ..] along axis d; one thread per output.
template <typename R>
__global__ void dense_axis_kernel(int d, int n0, int n1, int n2, const R *__restrict__ T, int trans,
                                  const cpx<R> *__restrict__ x, cpx<R> *__restrict__ y)
{
    const long N = (long)n0 * n1 * n2;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int i0 = (int)(I % n0), i1 = (int)((I / n0) % n1);
    const int i2 = (int)(I / ((long)n0 * n1));
    int n, i;
    long base, stride;
    if (d == 0) { n = n0; i = i0; stride = 1; base = I - i0; }
    else if (d == 1) { n = n1; i = i1; stride = n0; base = I - (long)i1 * n0; }
    else { n = n2; i = i2; stride = (long)n0 * n1; base = I - (long)i2 * n0 * n1; }
    R sr = 0, si = 0;
    for (int k = 0; k < n; ++k) {
        const R tv = trans ? T[(long)k * n + i] : T[(long)i * n + k];
        const cpx<R> v = x[base + k * stride];
        sr += tv * v.re; si += tv * v.im;
    }
    y[I] = {sr, si};
}

// Tiled form of the same transform as a batched complex x real GEMM:
//   Out[b][i][l] = sum_k T(i, k) In[b][k][l],  element (b, k, l) at b*sB + k*sK + l*sL.
// A 64 x 64 (i, l) output tile per workgroup, K in chunks of 16 staged in LDS,
// each of the 256 threads accumulates a 4 x 4 block. kfast: the input tile is
// read with k fastest (sK = 1), else l fastest (sL = 1).
constexpr int GT = 64, GK = 16;
template <typename R>
__global__ void __launch_bounds__(256) axis_gemm_kernel(int n, long L, long sB, long sK, long sL, int kfast,
                                                        const R *__restrict__ T, int trans,
                                                        const cpx<R> *__restrict__ in, cpx<R> *__restrict__ out)
{
    __shared__ R Ts[GK][GT + 1];      // [k][i]
    __shared__ cpx<R> Is[GK][GT + 1];    // [k][l]
    const int tid = threadIdx.x;
    const long l0 = (long)blockIdx.x * GT;
    const int i0 = blockIdx.y * GT;
    const long boff = (long)blockIdx.z * sB;
    const int ti = tid / 16, tl = tid % 16;   // 16 x 16 threads, 4 x 4 outputs each
    R ar[4][4], ai[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) ar[a][b] = ai[a][b] = 0.0;
    constexpr int NE = GK * GT / 256;
    // T tile Ts[k][i] = T(i0 + i, k0 + k) and the input tile; software pipeline: the
    // global loads of chunk k0 + GK are in flight while chunk k0 is multiplied out of LDS
    R tv[NE];
    cpx<R> uv[NE];
    auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            const int e = tid + 256 * r;
            const int kk = e % GK, ii = e / GK;
            const int gi = i0 + ii, gk = k0 + kk;
            tv[r] = 0.0;
            if (gi < n && gk < n) tv[r] = trans ? T[(long)gk * n + gi] : T[(long)gi * n + gk];
            int kq, lq;
            if (kfast) { kq = e % GK; lq = e / GK; } else { lq = e % GT; kq = e / GT; }
            const int gk2 = k0 + kq;
            const long gl = l0 + lq;
            uv[r] = {0.0, 0.0};
            if (gk2 < n && gl < L) uv[r] = in[boff + gk2 * sK + gl * sL];
        }
    };
    load(0);
    for (int k0 = 0; k0 < n; k0 += GK) {
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            const int e = tid + 256 * r;
            Ts[e % GK][e / GK] = tv[r];
            int kq, lq;
            if (kfast) { kq = e % GK; lq = e / GK; } else { lq = e % GT; kq = e / GT; }
            Is[kq][lq] = uv[r];
        }
        __syncthreads();
        if (k0 + GK < n) load(k0 + GK);
#pragma unroll 4
        for (int kk = 0; kk < GK; ++kk) {
            R tv[4];
            cpx<R> iv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) tv[a] = Ts[kk][ti + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) iv[b] = Is[kk][tl + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    ar[a][b] += tv[a] * iv[b].re;
                    ai[a][b] += tv[a] * iv[b].im;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int gi = i0 + ti + 16 * a;
        if (gi >= n) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const long gl = l0 + tl + 16 * b;
            if (gl < L) out[boff + gi * sK + gl * sL] = {ar[a][b], ai[a][b]};
        }
    }
}

// FP32 variant on MFMA (v_mfma_f32_16x16x4_f32; lane l holds A[l%16][l/16],
// B[l/16][l%16], D[4(l/16)+q][l%16]): the same 64 x 64 output tile per block, staged
// through LDS as above; wave (wi, wl) of the 2 x 2 waves owns a 32 x 32 quadrant as
// 2 x 2 MFMA tiles, real and imaginary parts as two real GEMMs with the same A = T.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) axis_gemm_mfma_kernel(int n, long L, long sB, long sK, long sL, int kfast,
                                                             const float *__restrict__ T, int trans,
                                                             const fcomplex *__restrict__ in,
                                                             fcomplex *__restrict__ out)
{
    __shared__ float Ts[GK][GT];      // [k][i]
    __shared__ float Ir[GK][GT], Ii[GK][GT];   // [k][l]
    const int tid = threadIdx.x;
    const long l0 = (long)blockIdx.x * GT;
    const int i0 = blockIdx.y * GT;
    const long boff = (long)blockIdx.z * sB;
    const int lane = tid & 63, wv = tid >> 6, wi = wv >> 1, wl = wv & 1;
    const int r16 = lane & 15, k4 = lane >> 4;
    f32x4 accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) accr[a][b] = acci[a][b] = {0.f, 0.f, 0.f, 0.f};
    constexpr int NE = GK * GT / 256;
    // software pipeline: the global loads of chunk k0 + GK are in flight while chunk k0
    // is multiplied out of LDS
    float tv[NE];
    fcomplex uv[NE];
    auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            const int e = tid + 256 * r;
            const int kk = e % GK, ii = e / GK;
            const int gi = i0 + ii, gk = k0 + kk;
            tv[r] = 0.f;
            if (gi < n && gk < n) tv[r] = trans ? T[(long)gk * n + gi] : T[(long)gi * n + gk];
            int kq, lq;
            if (kfast) { kq = e % GK; lq = e / GK; } else { lq = e % GT; kq = e / GT; }
            const int gk2 = k0 + kq;
            const long gl = l0 + lq;
            uv[r] = {0.f, 0.f};
            if (gk2 < n && gl < L) uv[r] = in[boff + gk2 * sK + gl * sL];
        }
    };
    load(0);
    for (int k0 = 0; k0 < n; k0 += GK) {
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            const int e = tid + 256 * r;
            Ts[e % GK][e / GK] = tv[r];
            int kq, lq;
            if (kfast) { kq = e % GK; lq = e / GK; } else { lq = e % GT; kq = e / GT; }
            Ir[kq][lq] = uv[r].re;
            Ii[kq][lq] = uv[r].im;
        }
        __syncthreads();
        if (k0 + GK < n) load(k0 + GK);
#pragma unroll
        for (int s = 0; s < GK / 4; ++s) {
            const int kk = 4 * s + k4;
            float av[2], br[2], bi[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = Ts[kk][32 * wi + 16 * a + r16];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                br[b] = Ir[kk][32 * wl + 16 * b + r16];
                bi[b] = Ii[kk][32 * wl + 16 * b + r16];
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], br[b], accr[a][b], 0, 0, 0);
                    acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bi[b], acci[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gi = i0 + 32 * wi + 16 * a + 4 * k4 + q;
                const long gl = l0 + 32 * wl + 16 * b + r16;
                if (gi < n && gl < L) out[boff + gi * sK + gl * sL] = {accr[a][b][q], acci[a][b][q]};
            }
}

// FP32 transform along one axis on MFMA (v_mfma_f32_16x16x4_f32; real and imaginary
// parts as two GEMMs sharing A), one wave per (16-column tile, group of G 16-row tiles)
// of the output, no LDS and no barriers. The matrix comes in fragment order (built on
// the host by axis_fragments): float4 s4 of lane l of row tile it holds
// A(16 it + l % 16, 4 (4 s4 + j) + l / 16), j < 4, zero past n, so a wave reads its rows
// with KS/4 fully coalesced 16-byte loads. The wave loads its 16 input columns as B
// fragments (KS k-steps of 4, branch-free: the index is clamped and the value zeroed
// past n), then runs the K = n reduction in 4G independent accumulator chains. Columns
// are numbered across the batch: column c starts at (c / Lin) sB + (c % Lin) sL, k
// steps by sK. Work item j = 4 t + wave (t the XCD-remapped block index) is column tile
// j / ngr, row group j % ngr: the row groups of one column tile run on one CU and
// neighbouring column tiles on one XCD, which share the input lines through L1 and L2.
struct DiagF {   // epilogue of the last forward pass: divide by alpha + g (l0 + l1 + l2) off the walls
    const float *lam[3];
    const unsigned char *wall[3];
    fcomplex alpha, g;
    int n0;
};

// FP64 transform along one axis on FP64 MFMA (v_mfma_f64_16x16x4_f64; lane l holds
// A[l%16][l/16], B[l/16][l%16] and D[4q+l/16][l%16], q < 4: rows interleaved across the lane
// groups, unlike the FP32 and FP16 forms). One wave per (16-column tile,
// 16-row tile) of the output: its KS k-steps of the matrix row tile (fragment order,
// double2 per lane: A(16 it + l % 16, 4 (2 s2 + j) + l / 16), zero past n) and its 16 input
// columns (branch-free: clamped index, zeroed value) in registers, the real and imaginary
// parts as two products sharing A, each in two accumulator chains. Column numbering and the
// XCD-aware work mapping are those of axis_wave_mfma_kernel (here the row tiles of a column
// tile are its "row groups": they run on one CU and share the input lines through L1).
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <int KS>
__global__ void __launch_bounds__(256) axis_wave_f64_kernel(int n, long ncol, long Lin, long sB, long sL, long sK,
                                                            int nit, const double2 *__restrict__ Tf,
                                                            const dcomplex *__restrict__ in,
                                                            dcomplex *__restrict__ out)
{
    constexpr int KS2 = KS / 2;
    const int nb = gridDim.x, b = blockIdx.x, per = nb / 8;
    const long t = b < 8 * per ? (long)(b % 8) * per + b / 8 : b;
    const long item = 4 * t + (threadIdx.x >> 6);
    const long nct = (ncol + 15) / 16;
    if (item >= nct * nit) return;   // whole wave: no barriers below
    const int lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const int it = (int)(item % nit);
    const long col = (item / nit) * 16 + r16;
    const bool cv = col < ncol;
    const long base = cv ? (col / Lin) * sB + (col % Lin) * sL : 0;
    double a[KS];
#pragma unroll
    for (int s2 = 0; s2 < KS2; ++s2) {
        const double2 v = Tf[((long)it * KS2 + s2) * 64 + lane];
        a[2 * s2] = v.x;
        a[2 * s2 + 1] = v.y;
    }
    dcomplex bv[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + k4;
        bv[s] = in[base + (long)min(k, n - 1) * sK];
        if (!cv || k >= n) bv[s] = {0.0, 0.0};
    }
    f64x4 ar[2], ai[2];
    ar[0] = ar[1] = ai[0] = ai[1] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        ar[s & 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], bv[s].re, ar[s & 1], 0, 0, 0);
        ai[s & 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], bv[s].im, ai[s & 1], 0, 0, 0);
    }
    if (!cv) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = 16 * it + 4 * q + k4;   // the FP64 form interleaves the rows
        if (i < n) out[base + (long)i * sK] = {ar[0][q] + ar[1][q], ai[0][q] + ai[1][q]};
    }
}

template <int KS, int G, bool DIAG>
__global__ void __launch_bounds__(256) axis_wave_mfma_kernel(int n, long ncol, long Lin, long sB, long sL, long sK,
                                                             int nit, const float4 *__restrict__ Tf,
                                                             const fcomplex *__restrict__ in,
                                                             fcomplex *__restrict__ out, DiagF dg)
{
    constexpr int KS4 = (KS + 3) / 4;
    const int ngr = (nit + G - 1) / G;
    const int nb = gridDim.x, b = blockIdx.x, per = nb / 8;
    const long t = b < 8 * per ? (long)(b % 8) * per + b / 8 : b;
    const long item = 4 * t + (threadIdx.x >> 6);
    const long nct = (ncol + 15) / 16;
    if (item >= nct * ngr) return;   // whole wave: no barriers below
    const int lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const int ig = (int)(item % ngr);
    const long col = (item / ngr) * 16 + r16;
    const bool cv = col < ncol;
    const long base = cv ? (col / Lin) * sB + (col % Lin) * sL : 0;
    float a[G][KS4 * 4];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int it = min(ig * G + g, nit - 1);   // past the last row tile: a copy, not stored
#pragma unroll
        for (int s4 = 0; s4 < KS4; ++s4) {
            const float4 v = Tf[(it * KS4 + s4) * 64 + lane];
            a[g][4 * s4] = v.x;
            a[g][4 * s4 + 1] = v.y;
            a[g][4 * s4 + 2] = v.z;
            a[g][4 * s4 + 3] = v.w;
        }
    }
    fcomplex bv[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + k4;
        bv[s] = in[base + (long)min(k, n - 1) * sK];
        if (!cv || k >= n) bv[s] = {0.f, 0.f};
    }
    f32x4 ar[G][2], ai[G][2];
#pragma unroll
    for (int g = 0; g < G; ++g) ar[g][0] = ar[g][1] = ai[g][0] = ai[g][1] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            ar[g][s & 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[g][s], bv[s].re, ar[g][s & 1], 0, 0, 0);
            ai[g][s & 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[g][s], bv[s].im, ai[g][s & 1], 0, 0, 0);
        }
    // lane holds D[4 k4 + q][r16] of each row tile: output row 16 it + 4 k4 + q of column col
    if (!cv) return;
    float la = 0.f;
    bool wl = false;
    if constexpr (DIAG) {   // z pass: column col = i0 + n0 i1, row i = i2
        const int i0 = (int)(col % dg.n0), i1 = (int)(col / dg.n0);
        la = dg.lam[0][i0] + dg.lam[1][i1];
        wl = (dg.wall[0][i0] | dg.wall[1][i1]) != 0;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int it = ig * G + g;
        if (it >= nit) break;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 16 * it + 4 * k4 + q;
            if (i >= n) continue;
            fcomplex v = {ar[g][0][q] + ar[g][1][q], ai[g][0][q] + ai[g][1][q]};
            if constexpr (DIAG) {
                if (!wl && !dg.wall[2][i]) {
                    const float lam = la + dg.lam[2][i];
                    const float dr = dg.alpha.re + dg.g.re * lam, di = dg.alpha.im + dg.g.im * lam;
                    const float inv = 1.f / (dr * dr + di * di);
                    v = {(v.re * dr + v.im * di) * inv, (v.im * dr - v.re * di) * inv};
                }
            }
            out[base + (long)i * sK] = v;
        }
    }
}

// The same transform in split FP16 on v_mfma_f32_16x16x32_f16 (lane l holds A[l%16][8(l/16)+j],
// B[8(l/16)+j][l%16], j < 8, and D[4(l/16)+q][l%16]). Each operand is a sum of two FP16
// numbers, hi = fp16(v) and lo = fp16(v - hi), and a product is hi*hi + hi*lo + lo*hi in FP32
// accumulation (the dropped lo*lo term and the splits leave about 2^-22 relative). The matrix
// comes pre-split and scaled by a power of two (|A| <= 1) in fragment order: uint4 (8 halves)
// [it][kb][hi, lo][lane]. The input column is scaled by a power of two from its largest
// component (a 4-lane maximum: the lanes l, l^16, l^32, l^48 hold one column), so that its
// values use FP16's range; both scales are undone exactly in the epilogue. Work mapping,
// column numbering and the diagonal epilogue are those of axis_wave_mfma_kernel.
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
struct fcomplex2 {
    fcomplex a, b;
};

// Loads and stores are raw buffer accesses with 32-bit byte offsets (the vectors hold at most
// 2^32 bytes; nbytes is their size): a lane's offset is its column start plus its k group,
// the (kb, j) step a wave-uniform stride, one integer add per access. Every offset is in
// bounds (k past the axis, only in the last k block, reads row 0 and is zeroed; padding
// columns past ncol read the last column, and their output columns are not stored); the
// buffer's own range check is a second guard.
// KF: the axis is the contiguous one (sK = 1, even line length): a lane's 8 consecutive k of a
// k block and its 4 consecutive output rows are 16-byte loads and stores (pairs of values).
// SPLIT false: one FP16 product per MFMA (hi halves only, about 2^-11 relative), for solves
// whose result enters as a small correction.
template <int KB, int G, bool DIAG, int OCC, bool KF, bool SPLIT = true>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC, 8))) axis_split16_kernel(int n, long ncol, long Lin, long sB, long sL, long sK,
                                                           int nit, const uint4 *__restrict__ Th, float ainv,
                                                           const fcomplex *__restrict__ in,
                                                           fcomplex *__restrict__ out, DiagF dg, unsigned nbytes)
{
    const int ngr = (nit + G - 1) / G;
    const int nb = gridDim.x, b = blockIdx.x, per = nb / 8;
    const long t = b < 8 * per ? (long)(b % 8) * per + b / 8 : b;
    const long item = 4 * t + (threadIdx.x >> 6);
    const long nct = (ncol + 15) / 16;
    if (item >= nct * ngr) return;   // whole wave: no barriers below
    const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
    const int ig = (int)(item % ngr);
    const long col = (item / ngr) * 16 + r16;
    const bool cv = col < ncol;
    const unsigned colu = (unsigned)(cv ? col : ncol - 1), linu = (unsigned)Lin, cq = colu / linu;
    const unsigned base8 = 8u * (cq * (unsigned)sB + (colu - cq * linu) * (unsigned)sL);   // column start, bytes
    const unsigned skb = 8u * (unsigned)sK;                                                  // bytes per k
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<fcomplex *>(in), (short)0, (int)nbytes, 0x00020000);
    const unsigned vin = base8 + (unsigned)(8 * kq) * skb;
    fcomplex bv[KB][8];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        if (KF && kb < KB - 1) {   // k < 32 (KB - 1) < n: four 16-byte loads of consecutive k
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const fcomplex2 v = __builtin_bit_cast(
                    fcomplex2, __builtin_amdgcn_raw_buffer_load_b128(rin, vin + (unsigned)(32 * kb + 2 * jj) * skb, 0, 0));
                bv[kb][2 * jj] = v.a;
                bv[kb][2 * jj + 1] = v.b;
            }
            continue;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 32 * kb + 8 * kq + j;
            // k < 32 (KB - 1) < n in every block but the last
            const bool in_axis = kb < KB - 1 || k < n;
            const unsigned off = in_axis ? vin + (unsigned)(32 * kb + j) * skb : base8;
            bv[kb][j] = __builtin_bit_cast(fcomplex, __builtin_amdgcn_raw_buffer_load_b64(rin, off, 0, 0));
            if (!in_axis) bv[kb][j] = {0.f, 0.f};
        }
    }
    float m = 0.f;
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) m = fmaxf(m, fmaxf(fabsf(bv[kb][j].re), fabsf(bv[kb][j].im)));
    m = fmaxf(m, __shfl_xor(m, 16, 64));
    m = fmaxf(m, __shfl_xor(m, 32, 64));
    int e = 0;
    if (m > 0.f) frexpf(m, &e);
    const float sc = ldexpf(1.f, -e);   // |v sc| < 1
    half8_t rh[KB], rl[KB], ih[KB], il[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float xr = bv[kb][j].re * sc, xi = bv[kb][j].im * sc;
            const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
            rh[kb][j] = hr;
            ih[kb][j] = hi;
            if (SPLIT) {
                rl[kb][j] = (_Float16)(xr - (float)hr);
                il[kb][j] = (_Float16)(xi - (float)hi);
            }
        }
    f32x4 ar[G], ai[G];
#pragma unroll
    for (int g = 0; g < G; ++g) ar[g] = ai[g] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int it = min(ig * G + g, nit - 1);   // past the last row tile: a copy, not stored
            const half8_t ah = __builtin_bit_cast(half8_t, Th[((long)(it * KB + kb) * 2) * 64 + lane]);
            if (SPLIT) {
                const half8_t al = __builtin_bit_cast(half8_t, Th[((long)(it * KB + kb) * 2 + 1) * 64 + lane]);
                ar[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, rh[kb], ar[g], 0, 0, 0);
                ar[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, rl[kb], ar[g], 0, 0, 0);
                ai[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, ih[kb], ai[g], 0, 0, 0);
                ai[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, il[kb], ai[g], 0, 0, 0);
            }
            ar[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, rh[kb], ar[g], 0, 0, 0);
            ai[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, ih[kb], ai[g], 0, 0, 0);
        }
    if (!cv) return;
    const float back = ainv / sc;   // powers of two: exact
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(out, (short)0, (int)nbytes, 0x00020000);
    const unsigned vout = base8 + (unsigned)(4 * kq) * skb;
    float la = 0.f;
    bool wl = false;
    if constexpr (DIAG) {   // z pass: column col = i0 + n0 i1, row i = i2
        const int i0 = (int)(col % dg.n0), i1 = (int)(col / dg.n0);
        la = dg.lam[0][i0] + dg.lam[1][i1];
        wl = (dg.wall[0][i0] | dg.wall[1][i1]) != 0;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int it = ig * G + g;
        if (it >= nit) break;
        if (KF && 16 * it + 4 * kq + 3 < n) {   // four consecutive rows: two 16-byte stores
#pragma unroll
            for (int qq = 0; qq < 2; ++qq) {
                const fcomplex2 v = {{ar[g][2 * qq] * back, ai[g][2 * qq] * back},
                                     {ar[g][2 * qq + 1] * back, ai[g][2 * qq + 1] * back}};
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), rout,
                                                       vout + (unsigned)(16 * it + 2 * qq) * skb, 0, 0);
            }
            continue;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 16 * it + 4 * kq + q;
            if (i >= n) continue;
            fcomplex v = {ar[g][q] * back, ai[g][q] * back};
            if constexpr (DIAG) {   // branch-free; the hardware reciprocal (about 1 ulp)
                const float lam = la + dg.lam[2][i];
                const float dr = dg.alpha.re + dg.g.re * lam, di = dg.alpha.im + dg.g.im * lam;
                const float inv = __builtin_amdgcn_rcpf(dr * dr + di * di);
                const fcomplex d = {(v.re * dr + v.im * di) * inv, (v.im * dr - v.re * di) * inv};
                if (!(wl || dg.wall[2][i])) v = d;
            }
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), rout,
                                                  vout + (unsigned)(16 * it + q) * skb, 0, 0);
        }
    }
}

// The split-FP16 transform with a wave per 16-column tile and ALL NIT row tiles: the input
// column tile is loaded, scaled and split once (axis_split16_kernel splits it once per
// group of three row tiles), and the matrix's fragments are staged once per block in LDS,
// where every wave of the block reads them (ds_read_b128, one 16-byte entry per lane).
// LDS layout per row tile: the KB - 1 full k blocks as [kb][hi, lo][64 lanes], then the last
// k block with only its first vg lane groups (the groups holding some k < n), [hi, lo][16 vg];
// a lane of a later group reads lane r16 of the first group instead, a finite value whose
// product with the zeroed input (k >= n) vanishes. At n = 132 that is 78 KB per block, so
// two blocks share a CU. Same operand layout, scaling, buffer accesses and epilogues as
// axis_split16_kernel.
template <int NIT, bool DIAG, bool KF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
axis_split16_lds_kernel(int n, long ncol, long Lin, long sB, long sL, long sK, const uint4 *__restrict__ Th, int vg,
                        float ainv, const fcomplex *__restrict__ in, fcomplex *__restrict__ out, DiagF dg,
                        unsigned nbytes)
{
    constexpr int KB = (NIT + 1) / 2;
    extern __shared__ uint4 As[];
    const int itS = (KB - 1) * 128 + 32 * vg;   // entries per row tile
    const int nb = gridDim.x, b = blockIdx.x, per = nb / 8;
    const long t = b < 8 * per ? (long)(b % 8) * per + b / 8 : b;
    const long ct = 4 * t + (threadIdx.x >> 6);
    const long nct = (ncol + 15) / 16;
    const bool live = ct < nct;   // a wave past the last tile only helps stage the matrix
    const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
    const long col = ct * 16 + r16;
    const bool cv = live && col < ncol;
    const unsigned colu = (unsigned)(cv ? col : ncol - 1), linu = (unsigned)Lin, cq = colu / linu;
    const unsigned base8 = 8u * (cq * (unsigned)sB + (colu - cq * linu) * (unsigned)sL);
    const unsigned skb = 8u * (unsigned)sK;
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<fcomplex *>(in), (short)0, (int)nbytes, 0x00020000);
    const unsigned vin = base8 + (unsigned)(8 * kq) * skb;
    fcomplex bv[KB][8];
    if (live) {   // the column loads go out before the matrix is staged
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            if (KF && kb < KB - 1) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const fcomplex2 v = __builtin_bit_cast(
                        fcomplex2,
                        __builtin_amdgcn_raw_buffer_load_b128(rin, vin + (unsigned)(32 * kb + 2 * jj) * skb, 0, 0));
                    bv[kb][2 * jj] = v.a;
                    bv[kb][2 * jj + 1] = v.b;
                }
                continue;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 32 * kb + 8 * kq + j;
                const bool in_axis = kb < KB - 1 || k < n;
                const unsigned off = in_axis ? vin + (unsigned)(32 * kb + j) * skb : base8;
                bv[kb][j] = __builtin_bit_cast(fcomplex, __builtin_amdgcn_raw_buffer_load_b64(rin, off, 0, 0));
                if (!in_axis) bv[kb][j] = {0.f, 0.f};
            }
        }
    }
    // stage the matrix (source: the padded fragment order [it][kb][hi, lo][64 lanes])
    const int full = (KB - 1) * 128, last = 16 * vg;
    for (int e = threadIdx.x; e < NIT * itS; e += 256) {
        const int it = e / itS, rem = e - it * itS;
        int src;
        if (rem < full) {
            src = it * KB * 128 + rem;
        } else {
            const int r2 = rem - full, h = r2 >= last ? 1 : 0, l = r2 - h * last;
            src = ((it * KB + KB - 1) * 2 + h) * 64 + l;
        }
        As[e] = Th[src];
    }
    __syncthreads();
    if (!live) return;
    float m = 0.f;
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) m = fmaxf(m, fmaxf(fabsf(bv[kb][j].re), fabsf(bv[kb][j].im)));
    m = fmaxf(m, __shfl_xor(m, 16, 64));
    m = fmaxf(m, __shfl_xor(m, 32, 64));
    int e = 0;
    if (m > 0.f) frexpf(m, &e);
    const float sc = ldexpf(1.f, -e);
    half8_t rh[KB], rl[KB], ih[KB], il[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float xr = bv[kb][j].re * sc, xi = bv[kb][j].im * sc;
            const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
            rh[kb][j] = hr;
            rl[kb][j] = (_Float16)(xr - (float)hr);
            ih[kb][j] = hi;
            il[kb][j] = (_Float16)(xi - (float)hi);
        }
    f32x4 ar[NIT], ai[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) ar[it] = ai[it] = {0.f, 0.f, 0.f, 0.f};
    const int ll = lane < last ? lane : r16;
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int o = it * itS + kb * 128;
            const bool lb = kb == KB - 1;
            const half8_t ah = __builtin_bit_cast(half8_t, As[o + (lb ? ll : lane)]);
            const half8_t al = __builtin_bit_cast(half8_t, As[o + (lb ? last + ll : 64 + lane)]);
            ar[it] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, rh[kb], ar[it], 0, 0, 0);
            ai[it] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, ih[kb], ai[it], 0, 0, 0);
            ar[it] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, rl[kb], ar[it], 0, 0, 0);
            ai[it] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, il[kb], ai[it], 0, 0, 0);
            ar[it] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, rh[kb], ar[it], 0, 0, 0);
            ai[it] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, ih[kb], ai[it], 0, 0, 0);
        }
    if (!cv) return;
    const float back = ainv / sc;
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(out, (short)0, (int)nbytes, 0x00020000);
    const unsigned vout = base8 + (unsigned)(4 * kq) * skb;
    float la = 0.f;
    bool wl = false;
    if constexpr (DIAG) {
        const int i0 = (int)(col % dg.n0), i1 = (int)(col / dg.n0);
        la = dg.lam[0][i0] + dg.lam[1][i1];
        wl = (dg.wall[0][i0] | dg.wall[1][i1]) != 0;
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        if (KF && 16 * it + 4 * kq + 3 < n) {
#pragma unroll
            for (int qq = 0; qq < 2; ++qq) {
                const fcomplex2 v = {{ar[it][2 * qq] * back, ai[it][2 * qq] * back},
                                     {ar[it][2 * qq + 1] * back, ai[it][2 * qq + 1] * back}};
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), rout,
                                                       vout + (unsigned)(16 * it + 2 * qq) * skb, 0, 0);
            }
            continue;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 16 * it + 4 * kq + q;
            if (i >= n) continue;
            fcomplex v = {ar[it][q] * back, ai[it][q] * back};
            if constexpr (DIAG) {
                const float lam = la + dg.lam[2][i];
                const float dr = dg.alpha.re + dg.g.re * lam, di = dg.alpha.im + dg.g.im * lam;
                const float inv = __builtin_amdgcn_rcpf(dr * dr + di * di);
                const fcomplex d = {(v.re * dr + v.im * di) * inv, (v.im * dr - v.re * di) * inv};
                if (!(wl || dg.wall[2][i])) v = d;
            }
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), rout,
                                                  vout + (unsigned)(16 * it + q) * skb, 0, 0);
        }
    }
}

// largest axis length the register-resident transform takes; longer axes use the
// LDS-tiled kernel
constexpr int kWaveAxisMax = 160;

__global__ void lincomb_kernel(LinComb lc, dcomplex *y, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        double re = 0, im = 0;
        for (int k = 0; k < lc.k; ++k) {
            const dcomplex v = lc.v[k][i], c = lc.c[k];
            re += c.re * v.re - c.im * v.im;
            im += c.re * v.im + c.im * v.re;
        }
        y[i] = {re, im};
    }
}

// out = the binomial extrapolation of order min(order, *step) from x (the current state)
// and h1, h2 (the two before): weights (1), (2, -1), (3, -3, 1). The order comes from the
// device step counter, so a captured step extrapolates only over states it has.
__global__ void extrap_kernel(const int *__restrict__ step, int order, const dcomplex *__restrict__ x,
                              const dcomplex *__restrict__ h1, const dcomplex *__restrict__ h2,
                              dcomplex *__restrict__ out, long N)
{
    const int o = min(order, *step);
    const double w0 = o == 0 ? 1.0 : (o == 1 ? 2.0 : 3.0), w1 = o == 0 ? 0.0 : (o == 1 ? -1.0 : -3.0),
                 w2 = o == 2 ? 1.0 : 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        const dcomplex a = x[i], b = h1[i], c = h2[i];
        out[i] = {w0 * a.re + w1 * b.re + w2 * c.re, w0 * a.im + w1 * b.im + w2 * c.im};
    }
}

// y1 = sum lc1; y2 = y1 + sum lc2; y3 = sum lc3 (the predictor's right-hand side, guess
// residual and guess in one pass)
__global__ void lincomb3_kernel(LinComb l1, LinComb l2, LinComb l3, dcomplex *y1, dcomplex *y2, dcomplex *y3, long N)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        double ar = 0, ai = 0, br = 0, bi = 0, cr = 0, ci = 0;
        for (int k = 0; k < l1.k; ++k) {
            const dcomplex v = l1.v[k][i], c = l1.c[k];
            ar += c.re * v.re - c.im * v.im;
            ai += c.re * v.im + c.im * v.re;
        }
        br = ar;
        bi = ai;
        for (int k = 0; k < l2.k; ++k) {
            const dcomplex v = l2.v[k][i], c = l2.c[k];
            br += c.re * v.re - c.im * v.im;
            bi += c.re * v.im + c.im * v.re;
        }
        for (int k = 0; k < l3.k; ++k) {
            const dcomplex v = l3.v[k][i], c = l3.c[k];
            cr += c.re * v.re - c.im * v.im;
            ci += c.re * v.im + c.im * v.re;
        }
        y1[i] = {ar, ai};
        y2[i] = {br, bi};
        y3[i] = {cr, ci};
    }
}

// lincomb3 with the block sums of |y2|^2 and |y1|^2: partial[b], partial[nb + b]
__global__ void __launch_bounds__(kRedThreads) lincomb3_norm_kernel(LinComb l1, LinComb l2, LinComb l3, dcomplex *y1,
                                                                    dcomplex *y2, dcomplex *y3, long N,
                                                                    double *__restrict__ partial)
{
    __shared__ double sh[kRedThreads / 64];
    double s1 = 0, s2 = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        double ar = 0, ai = 0, br = 0, bi = 0, cr = 0, ci = 0;
        for (int k = 0; k < l1.k; ++k) {
            const dcomplex v = l1.v[k][i], c = l1.c[k];
            ar += c.re * v.re - c.im * v.im;
            ai += c.re * v.im + c.im * v.re;
        }
        br = ar;
        bi = ai;
        for (int k = 0; k < l2.k; ++k) {
            const dcomplex v = l2.v[k][i], c = l2.c[k];
            br += c.re * v.re - c.im * v.im;
            bi += c.re * v.im + c.im * v.re;
        }
        for (int k = 0; k < l3.k; ++k) {
            const dcomplex v = l3.v[k][i], c = l3.c[k];
            cr += c.re * v.re - c.im * v.im;
            ci += c.re * v.im + c.im * v.re;
        }
        y1[i] = {ar, ai};
        y2[i] = {br, bi};
        y3[i] = {cr, ci};
        s1 += ar * ar + ai * ai;
        s2 += br * br + bi * bi;
    }
    s1 = block_sum(s1, sh);
    s2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s2;
        partial[gridDim.x + blockIdx.x] = s1;
    }
}

// Mass and z dipole of the same vector in one z-marching pass (FP64, equal orders P):
//   m = Mz My Mx x,  d = Xz My Mx x
// (real bands; the x and y passes are shared). Same tiling, plane prefetch and output
// rings as sep3_kernel.
template <int P>
__global__ void __launch_bounds__(256) mass_dip_kernel(int n0, int n1, int n2, int kchunk,
                                                       const double *__restrict__ Mx, const double *__restrict__ My,
                                                       const double *__restrict__ Mz, const double *__restrict__ Xz,
                                                       const dcomplex *__restrict__ x, dcomplex *__restrict__ m,
                                                       dcomplex *__restrict__ dz, dcomplex *__restrict__ xc)
{
    constexpr int R = 2 * P + 1, T = 16, H = T + 2 * P;
    __shared__ dcomplex Xs[H * H];
    __shared__ dcomplex Ts[H * T];
    const int tid = threadIdx.x, ti = tid % T, tj = tid / T;
    const int tilesx = (n0 + T - 1) / T, tilesy = (n1 + T - 1) / T;
    int b = blockIdx.x;
    const int bx = b % tilesx;
    b /= tilesx;
    const int by = b % tilesy, bz = b / tilesy;
    const int i0 = bx * T, j0 = by * T, k0 = bz * kchunk, k1 = min(k0 + kchunk, n2);
    const int i = i0 + ti, j = j0 + tj;
    const bool own = i < n0 && j < n1;
    const int ic = min(i, n0 - 1), jc = min(j, n1 - 1);
    double mx[R], my[R];
    dcomplex am[R], ad[R];
#pragma unroll
    for (int l = 0; l < R; ++l) {
        mx[l] = Mx[ic * R + l];
        my[l] = My[jc * R + l];
        am[l] = ad[l] = {0.0, 0.0};
    }
    const long plane = (long)n0 * n1;
    const int Kbeg = max(k0 - P, 0), Kend = min(k1 + P, n2);
    constexpr int NX = (H * H + 255) / 256;
    dcomplex xv[NX];
    auto load_plane = [&](int K) {
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int e = tid + 256 * q;
            const int ii = i0 - P + e % H, jj = j0 - P + e / H;
            xv[q] = {0.0, 0.0};
            if (e < H * H && ii >= 0 && ii < n0 && jj >= 0 && jj < n1) xv[q] = x[K * plane + (long)jj * n0 + ii];
        }
    };
    load_plane(Kbeg);
    for (int Kb = (Kbeg / R) * R; Kb < k1 + P; Kb += R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int K = Kb + r;
            if (K >= Kbeg && K < Kend) {   // block-uniform
#pragma unroll
                for (int q = 0; q < NX; ++q) {
                    const int e = tid + 256 * q;
                    if (e < H * H) Xs[e] = xv[q];
                }
                __syncthreads();
                if (K + 1 < Kend) load_plane(K + 1);
                if (xc && own && K >= k0 && K < k1) xc[K * plane + (long)j * n0 + i] = Xs[(tj + P) * H + ti + P];
                for (int e = tid; e < H * T; e += 256) {   // e % T == ti
                    const int jj = e / T;
                    double tr = 0, tim = 0;
#pragma unroll
                    for (int l = 0; l < R; ++l) {
                        const dcomplex v = Xs[jj * H + ti + l];
                        tr += mx[l] * v.re;
                        tim += mx[l] * v.im;
                    }
                    Ts[e] = {tr, tim};
                }
                __syncthreads();
                double ar = 0, ai = 0;
#pragma unroll
                for (int l = 0; l < R; ++l) {
                    const dcomplex tv = Ts[(tj + l) * T + ti];
                    ar += my[l] * tv.re;
                    ai += my[l] * tv.im;
                }
#pragma unroll
                for (int d = -P; d <= P; ++d) {
                    const int k = K + d;
                    if (k >= k0 && k < k1) {
                        const double mz = Mz[k * R + P - d], xz = Xz[k * R + P - d];
                        dcomplex &o = am[(r + d + P) % R], &q = ad[(r + d + P) % R];
                        o.re += mz * ar;
                        o.im += mz * ai;
                        q.re += xz * ar;
                        q.im += xz * ai;
                    }
                }
            }
            const int ko = K - P;
            if (ko >= k0 && ko < k1 && own) {
                m[ko * plane + (long)j * n0 + i] = am[r];
                dz[ko * plane + (long)j * n0 + i] = ad[r];
            }
            am[r] = ad[r] = {0.0, 0.0};
        }
    }
}

template <typename R>
__global__ void diag_scale_kernel(int n0, int n1, long N, const R *l0, const R *l1, const R *l2,
                                  const unsigned char *w0, const unsigned char *w1, const unsigned char *w2,
                                  cpx<R> alpha, cpx<R> g, const cpx<R> *__restrict__ x, cpx<R> *__restrict__ y)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int i = (int)(I % n0), j = (int)((I / n0) % n1);
    const long k = I / ((long)n0 * n1);
    const cpx<R> v = x[I];
    if (w0[i] | w1[j] | w2[k]) { y[I] = v; return; }
    const R lam = l0[i] + l1[j] + l2[k];
    const R dr = alpha.re + g.re * lam, di = alpha.im + g.im * lam;
    const R inv = R(1) / (dr * dr + di * di);
    y[I] = {(v.re * dr + v.im * di) * inv, (v.im * dr - v.re * di) * inv};
}

// --------------------------------------------------------------- sparse axis
template <typename R>
__global__ void csr_axis_kernel(int d, int a0, int a1, int a2, int m, const int *__restrict__ rowptr,
                                const int *__restrict__ col, const R *__restrict__ val,
                                const cpx<R> *__restrict__ in, cpx<R> *__restrict__ out, int accumulate)
{
    // output dims: axis d replaced by m
    const int o0 = d == 0 ? m : a0, o1 = d == 1 ? m : a1, o2 = d == 2 ? m : a2;
    const long N = (long)o0 * o1 * o2;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int i0 = (int)(I % o0), i1 = (int)((I / o0) % o1);
    const int i2 = (int)(I / ((long)o0 * o1));
    const int r = d == 0 ? i0 : (d == 1 ? i1 : i2);
    R sr = 0, si = 0;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
        const int c = col[e];
        const int j0 = d == 0 ? c : i0, j1 = d == 1 ? c : i1, j2 = d == 2 ? c : i2;
        const cpx<R> v = in[j0 + (long)a0 * (j1 + (long)a1 * j2)];
        sr += val[e] * v.re; si += val[e] * v.im;
    }
    if (accumulate) { sr += out[I].re; si += out[I].im; }
    out[I] = {sr, si};
}

inline unsigned grid_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
inline unsigned grid_capped(long n, int bs, unsigned cap)
{
    const unsigned g = grid_for(n, bs);
    return g < cap ? g : cap;
}

}  // namespace

// ===================================================================== launchers
template <typename R>
void launch_pot_apply_generic(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                              cpx<R> *y)
{
    const int n1 = ax[0].p + 1, n2 = ax[1].p + 1, n3 = ax[2].p + 1;
    const int Q1 = ax[0].Q, Q2 = ax[1].Q, Q3 = ax[2].Q;
    size_t nr = (size_t)(n1 * Q1 + n2 * Q2 + n3 * Q3 + 2 * (Q1 + Q2 + Q3));
    nr = (nr + 1) & ~(size_t)1;
    const size_t nc = (size_t)n1 * n2 * n3 + (size_t)n3 * n2 * Q1 + (size_t)n3 * Q2 * Q1 + (size_t)Q3 * Q2 * Q1;
    const size_t bytes = nr * sizeof(R) + nc * sizeof(cpx<R>);
    const long nel = (long)ax[0].E * ax[1].E * ax[2].E;
    hipLaunchKernelGGL(pot_apply_generic_kernel<R>, dim3((unsigned)nel), dim3(128), bytes, st, ax[0], ax[1], ax[2],
                       pot, gr, gi, x, y);
}

template <typename R, int P, int Q, int TE, int CPT>
static void launch_pot3d(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                         cpx<R> *y, cpx<R> *scr)
{
    constexpr int NT = (TE * Q) * (TE * Q) / CPT, TW = TE + P;
    const int nbx = (ax[0].E + TE - 1) / TE, nby = (ax[1].E + TE - 1) / TE;
    const dim3 grid((unsigned)nbx, (unsigned)nby);
    if (pot.kind == 2)
        hipLaunchKernelGGL((pot3d_kernel<R, P, Q, TE, CPT, 2>), grid, dim3(NT), 0, st, ax[0], ax[1], ax[2], pot, gr,
                           gi, x, scr);
    else
        hipLaunchKernelGGL((pot3d_kernel<R, P, Q, TE, CPT, 1>), grid, dim3(NT), 0, st, ax[0], ax[1], ax[2], pot, gr,
                           gi, x, scr);
    const long N = (long)ax[0].n * ax[1].n * ax[2].n;
    hipLaunchKernelGGL(pot3d_sum_kernel<R>, dim3((unsigned)std::min<long>((N + 255) / 256, 1L << 20)), dim3(256), 0,
                       st, ax[0].n, ax[1].n, ax[2].n, TE, TW, nbx, nby, scr, y);
}

long pot3d_scratch_len(long N) { return 9 * N; }   // MW^2 <= 9 classes for every variant (TW <= 3 TE)

template <typename R, int P, int Q, int TXE, int TYE>
static void launch_pot2d(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                         cpx<R> *y)
{
    const dim3 grid((unsigned)((ax[0].E + TXE - 1) / TXE), (unsigned)((ax[1].E + TYE - 1) / TYE));
    constexpr int NT = TXE * Q * TYE * Q;
    if (pot.kind == 2)
        hipLaunchKernelGGL((pot2d_kernel<R, P, Q, TXE, TYE, 2>), grid, dim3(NT), 0, st, ax[0], ax[1], pot, gr, gi, x,
                           y);
    else
        hipLaunchKernelGGL((pot2d_kernel<R, P, Q, TXE, TYE, 1>), grid, dim3(NT), 0, st, ax[0], ax[1], pot, gr, gi, x,
                           y);
}

// tile shapes (elements) of the 2D kernel: 1x1 (64 threads), 2x1, 2x2, 4x2, 4x4 (1024)
constexpr int kPot2dVariants = 5;
int pot2d_variants(int p, int Q) { return (Q == 8 && p >= 3 && p <= 5) ? kPot2dVariants : 0; }

template <typename R, int P>
static void launch_pot2d_p(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                           cpx<R> *y, int variant)
{
    switch (variant) {
    case 0: launch_pot2d<R, P, 8, 1, 1>(st, ax, pot, gr, gi, x, y); break;
    case 1: launch_pot2d<R, P, 8, 2, 1>(st, ax, pot, gr, gi, x, y); break;
    case 2: launch_pot2d<R, P, 8, 2, 2>(st, ax, pot, gr, gi, x, y); break;
    case 3: launch_pot2d<R, P, 8, 4, 2>(st, ax, pot, gr, gi, x, y); break;
    default: launch_pot2d<R, P, 8, 4, 4>(st, ax, pot, gr, gi, x, y); break;
    }
}

template <typename R>
bool launch_pot_apply_2d(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                         cpx<R> *y, int variant)
{
    const int p = ax[0].p, Q = ax[0].Q;
    if (ax[1].p != p || ax[1].Q != Q) return false;
    if (pot.kind != 1 && pot.kind != 2) return false;
    if (variant < 0 || variant >= pot2d_variants(p, Q)) return false;
    if (p == 3) launch_pot2d_p<R, 3>(st, ax, pot, gr, gi, x, y, variant);
    else if (p == 4) launch_pot2d_p<R, 4>(st, ax, pot, gr, gi, x, y, variant);
    else launch_pot2d_p<R, 5>(st, ax, pot, gr, gi, x, y, variant);
    return true;
}

int pot3d_variants(int p, int Q)
{
    if (p == 4 && Q == 8) return 4;
    if (p == 3 && Q == 8) return 4;
    if (p == 3 && Q == 5) return 1;
    return 0;
}

template <typename R>
bool launch_pot_apply_3d(hipStream_t st, const AxisDevT<R> ax[3], PotParams pot, R gr, R gi, const cpx<R> *x,
                         cpx<R> *y, int variant, cpx<R> *scr)
{
    if (!scr) return false;
    const int p = ax[0].p, Q = ax[0].Q;
    for (int d = 1; d < 3; ++d)
        if (ax[d].p != p || ax[d].Q != Q) return false;
    if (pot.kind != 1 && pot.kind != 2) return false;
    if (variant < 0 || variant >= pot3d_variants(p, Q)) return false;
    // variant 0: 4 x 4 element tiles, 1024 threads; 1: 2 x 2 tiles, 256 threads; 2: 2 x 2
    // tiles, 128 threads of 2 point columns; 3: 4 x 4 tiles, 256 threads of 4 columns
    if (p == 4 && Q == 8) {
        switch (variant) {
        case 0: launch_pot3d<R, 4, 8, 4, 1>(st, ax, pot, gr, gi, x, y, scr); break;
        case 1: launch_pot3d<R, 4, 8, 2, 1>(st, ax, pot, gr, gi, x, y, scr); break;
        case 2: launch_pot3d<R, 4, 8, 2, 2>(st, ax, pot, gr, gi, x, y, scr); break;
        default: launch_pot3d<R, 4, 8, 4, 4>(st, ax, pot, gr, gi, x, y, scr); break;
        }
        return true;
    }
    if (p == 3 && Q == 8) {
        switch (variant) {
        case 0: launch_pot3d<R, 3, 8, 4, 1>(st, ax, pot, gr, gi, x, y, scr); break;
        case 1: launch_pot3d<R, 3, 8, 2, 1>(st, ax, pot, gr, gi, x, y, scr); break;
        case 2: launch_pot3d<R, 3, 8, 2, 2>(st, ax, pot, gr, gi, x, y, scr); break;
        default: launch_pot3d<R, 3, 8, 4, 4>(st, ax, pot, gr, gi, x, y, scr); break;
        }
        return true;
    }
    if (p == 3 && Q == 5) { launch_pot3d<R, 3, 5, 4, 1>(st, ax, pot, gr, gi, x, y, scr); return true; }
    return false;
}

template <typename R>
void launch_sep_apply(hipStream_t st, const SepBandsT<R> &sb, const cpx<R> *x, cpx<R> *y, cpx<R> *tmp[4])
{
    const int n0 = sb.n[0], n1 = sb.n[1], n2 = sb.n[2];
    const long N = (long)n0 * n1 * n2;
    const int bs = 256;
    if constexpr (std::is_same<R, double>::value) {
        const int p = sb.p[0];
        if (p == sb.p[1] && p == sb.p[2] && p >= 2 && p <= 5 && n2 > 2 * p) {
            // z chunks: one round of blocks at two per CU, each chunk at least 2P + 2 planes
            static int ncu = 0;
            if (!ncu) {
                int dev = 0;
                if (hipGetDevice(&dev) != hipSuccess ||
                    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
                    ncu = 256;
            }
            const int tiles = ((n0 + 15) / 16) * ((n1 + 15) / 16);
            const int chunks = std::max(1, std::min(2 * ncu / tiles, n2 / (2 * p + 2)));
            const int kchunk = (n2 + chunks - 1) / chunks;
            const unsigned nb = (unsigned)(tiles * ((n2 + kchunk - 1) / kchunk));
            const double *Mz = sb.Mz2 ? sb.Mz2 : sb.M[2];
#define TDZ_SEP3(PP)                                                                                             \
    hipLaunchKernelGGL(sep3_kernel<PP>, dim3(nb), dim3(256), 0, st, n0, n1, n2, kchunk, sb.M[0], sb.B[0], sb.M[1], \
                       sb.B[1], Mz, sb.B[2], x, y)
            switch (p) {
            case 2: TDZ_SEP3(2); break;
            case 3: TDZ_SEP3(3); break;
            case 4: TDZ_SEP3(4); break;
            default: TDZ_SEP3(5); break;
            }
#undef TDZ_SEP3
            return;
        }
    }
    hipLaunchKernelGGL(sep_x_kernel<R>, dim3(grid_for(N, bs)), dim3(bs), 0, st, n0, N / n0, sb.p[0], sb.M[0], sb.B[0],
                       x, tmp[0], tmp[1]);
    hipLaunchKernelGGL(sep_y_kernel<R>, dim3(grid_for(N, bs)), dim3(bs), 0, st, n0, n1, N, sb.p[1], sb.M[1], sb.B[1],
                       tmp[0], tmp[1], tmp[2], tmp[3]);
    hipLaunchKernelGGL(sep_z_kernel<R>, dim3(grid_for(N, bs)), dim3(bs), 0, st, (long)n0 * n1, n2, sb.p[2],
                       sb.Mz2 ? sb.Mz2 : sb.M[2], sb.B[2], tmp[2], tmp[3], y);
}

template <typename R>
void launch_wall_fix(hipStream_t st, const int n[3], const unsigned char *wall[3], const cpx<R> *x, cpx<R> *y)
{
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(wall_fix_kernel<R>, dim3(grid_for(N, 256)), dim3(256), 0, st, n[0], n[1], N, wall[0], wall[1],
                       wall[2], x, y);
}

void launch_multi_dot(hipStream_t st, const dcomplex *V, long ld, int nv, const dcomplex *w, long N,
                      dcomplex *partial, dcomplex *out)
{
    const unsigned nb = grid_capped(N, kRedThreads, kRedBlocks);
    for (int i0 = 0; i0 < nv; i0 += 4)
        hipLaunchKernelGGL(multi_dot_kernel<4>, dim3(nb), dim3(kRedThreads), 0, st, V, ld, i0, nv, w, N, partial);
    hipLaunchKernelGGL(reduce_complex_kernel, dim3(nv), dim3(kRedThreads), 0, st, partial, (int)nb, out);
}

void launch_multi_axpy_norm(hipStream_t st, const dcomplex *V, long ld, int nv, const dcomplex *h, dcomplex *w,
                            long N, double *partial, double *out_norm2)
{
    const unsigned nb = grid_capped(N, kRedThreads, kRedBlocks);
    hipLaunchKernelGGL(multi_axpy_norm_kernel, dim3(nb), dim3(kRedThreads), 0, st, V, ld, nv, h, w, N, partial);
    hipLaunchKernelGGL(reduce_real_kernel, dim3(1), dim3(kRedThreads), 0, st, partial, (int)nb, out_norm2);
}

void launch_multi_axpy_norm_dot(hipStream_t st, const dcomplex *V, long ld, int nv, const dcomplex *h, dcomplex *w,
                                long N, double *pnorm, dcomplex *pdot, double *out_norm2, dcomplex *out_dot)
{
    if (nv < 1 || nv > kCgsFusedMax) throw std::runtime_error("multi_axpy_norm_dot: 1 .. kCgsFusedMax vectors");
    const unsigned nb = grid_capped(N, kRedThreads, kRedBlocks);
    hipLaunchKernelGGL(multi_axpy_norm_dot_kernel<kCgsFusedMax>, dim3(nb), dim3(kRedThreads), 0, st, V, ld, nv, h, w,
                       N, pnorm, pdot);
    hipLaunchKernelGGL(reduce_real_kernel, dim3(1), dim3(kRedThreads), 0, st, pnorm, (int)nb, out_norm2);
    hipLaunchKernelGGL(reduce_complex_kernel, dim3(nv), dim3(kRedThreads), 0, st, pdot, (int)nb, out_dot);
}

void launch_correct_scale(hipStream_t st, const dcomplex *V, long ld, int nv, const dcomplex *h, dcomplex *w, long N,
                          double s)
{
    hipLaunchKernelGGL(correct_scale_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, V, ld, nv, h, w, N, s);
}

void launch_multi_axpy(hipStream_t st, const dcomplex *Z, long ld, int nv, const dcomplex *c, dcomplex *x, long N,
                       const dcomplex *base)
{
    hipLaunchKernelGGL(multi_axpy_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, Z, ld, nv, c,
                       base ? base : x, x, N, CoefArg{});
}

void launch_multi_axpy_args(hipStream_t st, const dcomplex *Z, long ld, int nv, const dcomplex *hc, dcomplex *x,
                            long N, const dcomplex *base)
{
    CoefArg cv{};
    for (int i = 0; i < nv && i < kCoefArgMax; ++i) cv.v[i] = hc[i];
    hipLaunchKernelGGL(multi_axpy_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, Z, ld, nv,
                       (const dcomplex *)nullptr, base ? base : x, x, N, cv);
}

// one thread per band entry, blockIdx.y the axis
__global__ void band_update_kernel(BandUpdate bu)
{
#pragma clang fp contract(off)
    const int d = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= bu.sz[d]) return;
    const dcomplex g = bu.g;
    // T = S / 2m + (-i W) + sE X, or 0 on an inactive axis
    double tr = 0.0, ti = 0.0, fr = 0.0, fi = 0.0;
    if (bu.active[d]) {
        const double w = bu.cap ? bu.W[d][k] : 0.0, sx = bu.sE[d] * bu.X[d][k];
        tr = bu.inv2m * bu.S[d][k] + 0.0 + sx;
        ti = -w;
        fr = 0.0 + sx;
        fi = -w;
    }
    dcomplex b = {g.re * tr - g.im * ti, g.re * ti + g.im * tr};
    if (d == 0) {
        const double m = bu.M[0][k];
        b.re += bu.alpha.re * m;
        b.im += bu.alpha.im * m;
    }
    bu.B[d][k] = b;
    if (bu.Bc[d]) bu.Bc[d][k] = {(float)(g.re * fr - g.im * fi), (float)(g.re * fi + g.im * fr)};
    if (bu.Bf[d]) bu.Bf[d][k] = {(float)b.re, (float)b.im};
}

void launch_band_update(hipStream_t st, const BandUpdate &bu)
{
    const int mx = std::max(bu.sz[0], std::max(bu.sz[1], bu.sz[2]));
    if (mx <= 0) return;
    hipLaunchKernelGGL(band_update_kernel, dim3((unsigned)((mx + 255) / 256), 3), dim3(256), 0, st, bu);
}

void launch_lincomb(hipStream_t st, const LinComb &lc, dcomplex *y, long N)
{
    if (lc.k < 1 || lc.k > LinComb::kMax) throw std::runtime_error("lincomb: 1 .. kMax terms");
    hipLaunchKernelGGL(lincomb_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, lc, y, N);
}

void launch_extrapolate(hipStream_t st, const int *step, int order, const dcomplex *x, const dcomplex *h1,
                        const dcomplex *h2, dcomplex *out, long N)
{
    hipLaunchKernelGGL(extrap_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, step, order, x, h1, h2, out,
                       N);
}

void launch_lincomb3(hipStream_t st, const LinComb &l1, const LinComb &l2, const LinComb &l3, dcomplex *y1,
                     dcomplex *y2, dcomplex *y3, long N)
{
    hipLaunchKernelGGL(lincomb3_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, l1, l2, l3, y1, y2, y3, N);
}

void launch_lincomb3_norms(hipStream_t st, const LinComb &l1, const LinComb &l2, const LinComb &l3, dcomplex *y1,
                           dcomplex *y2, dcomplex *y3, long N, double *partial, int nbmax, double *out)
{
    const unsigned nb = std::min(grid_capped(N, kRedThreads, 4096), (unsigned)nbmax);
    hipLaunchKernelGGL(lincomb3_norm_kernel, dim3(nb), dim3(kRedThreads), 0, st, l1, l2, l3, y1, y2, y3, N, partial);
    hipLaunchKernelGGL(reduce_real_kernel, dim3(2), dim3(kRedThreads), 0, st, partial, (int)nb, out);   // out[0], out[1]
}

bool launch_mass_dipole_z(hipStream_t st, const int n[3], int p, const double *M[3], const double *Xz,
                          const dcomplex *x, dcomplex *m, dcomplex *dz, dcomplex *xc)
{
    const int n0 = n[0], n1 = n[1], n2 = n[2];
    if (p < 2 || p > 5 || n2 <= 2 * p) return false;
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
    }
    // z chunks: one round of blocks at the kernel's occupancy (from the runtime, per order)
    static int per[6] = {};
    const int tiles = ((n0 + 15) / 16) * ((n1 + 15) / 16);
    int kchunk = 0;
    unsigned nb = 0;
    auto grid = [&](const void *fn) {
        if (!per[p]) {
            int b = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, 256, 0) != hipSuccess || b < 1) b = 2;
            (void)hipGetLastError();
            per[p] = b;
        }
        const int chunks = std::max(1, std::min(per[p] * ncu / tiles, n2 / (2 * p + 2)));
        kchunk = (n2 + chunks - 1) / chunks;
        nb = (unsigned)(tiles * ((n2 + kchunk - 1) / kchunk));
    };
#define TDZ_MD(PP)                                                                                              \
    grid(reinterpret_cast<const void *>(&mass_dip_kernel<PP>));                                                   \
    hipLaunchKernelGGL(mass_dip_kernel<PP>, dim3(nb), dim3(256), 0, st, n0, n1, n2, kchunk, M[0], M[1], M[2], Xz, x, \
                       m, dz, xc)
    switch (p) {
    case 2: TDZ_MD(2); break;
    case 3: TDZ_MD(3); break;
    case 4: TDZ_MD(4); break;
    default: TDZ_MD(5); break;
    }
#undef TDZ_MD
    return true;
}

void launch_axpby(hipStream_t st, dcomplex a, const dcomplex *x, dcomplex b, dcomplex *y, long N)
{
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, a, x, b, y, N);
}

template <typename R>
void launch_axpby_t(hipStream_t st, cpx<R> a, const cpx<R> *x, cpx<R> b, cpx<R> *y, long N)
{
    hipLaunchKernelGGL(axpby_t_kernel<R>, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, a, x, b, y, N);
}

void launch_demote(hipStream_t st, const dcomplex *x, fcomplex *y, long N)
{
    hipLaunchKernelGGL(demote_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, x, y, N);
}

// z (complex128) = x + y (complex64 sum, then widened)
__global__ void promote_add_kernel(const fcomplex *__restrict__ x, const fcomplex *__restrict__ y,
                                   dcomplex *__restrict__ z, long N)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < N; I += stride) {
        const fcomplex a = x[I], b = y[I];
        z[I] = {(double)(a.re + b.re), (double)(a.im + b.im)};
    }
}

void launch_promote_add(hipStream_t st, const fcomplex *x, const fcomplex *y, dcomplex *z, long N)
{
    hipLaunchKernelGGL(promote_add_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, x, y, z, N);
}

void launch_promote(hipStream_t st, const fcomplex *x, dcomplex *y, long N)
{
    hipLaunchKernelGGL(promote_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, x, y, N);
}

void launch_corr_combine(hipStream_t st, const fcomplex *r, const fcomplex *My, const fcomplex *c,
                         const fcomplex *n, const float *Vg, fcomplex g, fcomplex *q, long N, const float *no,
                         fcomplex ng)
{
    hipLaunchKernelGGL(corr_combine_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, r, My, c, n, Vg, g, q,
                       N, no, ng);
}

void launch_scale_invsqrt(hipStream_t st, const dcomplex *x, const double *norm2, dcomplex *y, long N)
{
    hipLaunchKernelGGL(scale_invsqrt_kernel, dim3(grid_capped(N, 256, 4096)), dim3(256), 0, st, x, norm2, y, N);
}

void launch_norm2(hipStream_t st, const dcomplex *x, long N, double *partial, double *out)
{
    const unsigned nb = grid_capped(N, kRedThreads, kRedBlocks);
    hipLaunchKernelGGL(norm2_kernel, dim3(nb), dim3(kRedThreads), 0, st, x, N, partial);
    hipLaunchKernelGGL(reduce_real_kernel, dim3(1), dim3(kRedThreads), 0, st, partial, (int)nb, out);
}

int axis_frag_steps(int n)
{
    const int ns = (n + 3) / 4;
    if (n > kWaveAxisMax) return 0;
    return ns <= 8 ? 8 : ns <= 16 ? 16 : ns <= 24 ? 24 : ns <= 33 ? 33 : 40;
}

std::vector<float> axis_fragments(int n, const std::vector<double> &T, bool trans)
{
    const int KS = axis_frag_steps(n);
    if (!KS) return {};
    const int KS4 = (KS + 3) / 4, nit = (n + 15) / 16;
    std::vector<float> f((size_t)nit * KS4 * 256, 0.f);
    for (int it = 0; it < nit; ++it)
        for (int s4 = 0; s4 < KS4; ++s4)
            for (int l = 0; l < 64; ++l)
                for (int j = 0; j < 4; ++j) {
                    const int row = 16 * it + l % 16, k = 4 * (4 * s4 + j) + l / 16;
                    if (row < n && k < n)
                        f[(((size_t)it * KS4 + s4) * 64 + l) * 4 + j] =
                            (float)(trans ? T[(size_t)k * n + row] : T[(size_t)row * n + k]);
                }
    return f;
}

int axis_f64_steps(int n) { return n > kWaveAxisMax ? 0 : 2 * ((n + 7) / 8); }

std::vector<double> axis_fragments_f64(int n, const std::vector<double> &T, bool trans)
{
    const int KS = axis_f64_steps(n);
    if (!KS) return {};
    const int KS2 = KS / 2, nit = (n + 15) / 16;
    std::vector<double> f((size_t)nit * KS2 * 128, 0.0);
    for (int it = 0; it < nit; ++it)
        for (int s2 = 0; s2 < KS2; ++s2)
            for (int l = 0; l < 64; ++l)
                for (int j = 0; j < 2; ++j) {
                    const int row = 16 * it + l % 16, k = 4 * (2 * s2 + j) + l / 16;
                    if (row < n && k < n)
                        f[(((size_t)it * KS2 + s2) * 64 + l) * 2 + j] = trans ? T[(size_t)k * n + row] : T[(size_t)row * n + k];
                }
    return f;
}

int axis_split16_blocks(int n) { return n > kWaveAxisMax ? 0 : (n + 31) / 32; }

bool split16_fits(const int n[3]) { return 8L * n[0] * n[1] * n[2] <= 0xFFFFFFFFL; }

std::vector<unsigned short> axis_fragments_split16(int n, const std::vector<double> &T, bool trans, float &ainv)
{
    const int KB = axis_split16_blocks(n);
    if (!KB) return {};
    double amax = 0;
    for (double v : T) amax = std::max(amax, std::fabs(v));
    int e = 0;
    if (amax > 0) std::frexp(amax, &e);
    const double as = std::ldexp(1.0, -e);   // |A as| < 1
    ainv = (float)std::ldexp(1.0, e);
    const int nit = (n + 15) / 16;
    std::vector<unsigned short> f((size_t)nit * KB * 2 * 64 * 8, 0);
    for (int it = 0; it < nit; ++it)
        for (int kb = 0; kb < KB; ++kb)
            for (int l = 0; l < 64; ++l)
                for (int j = 0; j < 8; ++j) {
                    const int row = 16 * it + l % 16, k = 32 * kb + 8 * (l / 16) + j;
                    if (row >= n || k >= n) continue;
                    const double v = as * (trans ? T[(size_t)k * n + row] : T[(size_t)row * n + k]);
                    const _Float16 hi = (_Float16)v, lo = (_Float16)(v - (double)hi);
                    const size_t o = ((((size_t)it * KB + kb) * 2) * 64 + l) * 8 + j;
                    f[o] = __builtin_bit_cast(unsigned short, hi);
                    f[o + 64 * 8] = __builtin_bit_cast(unsigned short, lo);
                }
    return f;
}

namespace {
// one launch of the LDS-staged form (a wave per column tile, all row tiles); false if the
// kernel cannot be given its LDS
template <int NIT, bool DIAG, bool KF>
bool launch_split16_lds_nit(hipStream_t st, int nd, long ncol, long Lin, long sB, long sL, long sK, const uint4 *T4,
                            float ainv, const fcomplex *x, fcomplex *y, const DiagF &dg, unsigned nbytes)
{
    constexpr int KB = (NIT + 1) / 2;
    const int vg = (nd - 32 * (KB - 1) + 7) / 8;
    const unsigned lds = (unsigned)(NIT * ((KB - 1) * 128 + 32 * vg) * 16);
    static int ok = -1;   // per instantiation: the dynamic LDS limit raised once
    if (ok < 0) {
        const void *fn = reinterpret_cast<const void *>(&axis_split16_lds_kernel<NIT, DIAG, KF>);
        ok = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess ? 1 : 0;
        (void)hipGetLastError();
    }
    if (!ok) return false;
    const unsigned nblk = (unsigned)(((ncol + 15) / 16 + 3) / 4);
    hipLaunchKernelGGL((axis_split16_lds_kernel<NIT, DIAG, KF>), dim3(nblk), dim3(256), lds, st, nd, ncol, Lin, sB,
                       sL, sK, T4, vg, ainv, x, y, dg, nbytes);
    return true;
}

template <bool DIAG, bool KF>
bool launch_split16_lds(hipStream_t st, int nd, long ncol, long Lin, long sB, long sL, long sK, const uint4 *T4,
                        float ainv, const fcomplex *x, fcomplex *y, const DiagF &dg, unsigned nbytes)
{
#define TDZ_SL(NN) return launch_split16_lds_nit<NN, DIAG, KF>(st, nd, ncol, Lin, sB, sL, sK, T4, ainv, x, y, dg, nbytes)
    switch ((nd + 15) / 16) {
    case 1: TDZ_SL(1);
    case 2: TDZ_SL(2);
    case 3: TDZ_SL(3);
    case 4: TDZ_SL(4);
    case 5: TDZ_SL(5);
    case 6: TDZ_SL(6);
    case 7: TDZ_SL(7);
    case 8: TDZ_SL(8);
    case 9: TDZ_SL(9);
    case 10: TDZ_SL(10);
    default: return false;
    }
#undef TDZ_SL
}

// split16 selects the form: 1 a wave per (column tile, group of three row tiles), matrix
// fragments through the caches; 2 the LDS-staged form
template <bool DIAG>
void launch_split16_axis(hipStream_t st, int nd, long ncol, long Lin, long sB, long sL, long sK, const unsigned short *Th,
                         float ainv, const fcomplex *x, fcomplex *y, const DiagF &dg, unsigned nbytes, int split16)
{
    if (split16 == 3) {   // single FP16 products
        const int nit = (nd + 15) / 16, G = nit < 3 ? nit : 3, ngr = (nit + G - 1) / G;
        const unsigned nblk = (unsigned)(((ncol + 15) / 16 * ngr + 3) / 4);
        const uint4 *T4 = reinterpret_cast<const uint4 *>(Th);
        const bool kf = !DIAG && sK == 1 && nd % 2 == 0;
#define TDZ_S16L(KB, GG)                                                                                            \
    do {                                                                                                        \
        if (kf)                                                                                                 \
            hipLaunchKernelGGL((axis_split16_kernel<KB, GG, false, 4, true, false>), dim3(nblk), dim3(256), 0, st, \
                               nd, ncol, Lin, sB, sL, sK, nit, T4, ainv, x, y, dg, nbytes);                     \
        else                                                                                                    \
            hipLaunchKernelGGL((axis_split16_kernel<KB, GG, DIAG, 4, false, false>), dim3(nblk), dim3(256), 0, st, \
                               nd, ncol, Lin, sB, sL, sK, nit, T4, ainv, x, y, dg, nbytes);                     \
    } while (0)
#define TDZ_S16L_G(KB)                 \
    do {                               \
        if (G == 1) TDZ_S16L(KB, 1);   \
        else if (G == 2) TDZ_S16L(KB, 2); \
        else TDZ_S16L(KB, 3);          \
    } while (0)
        switch (axis_split16_blocks(nd)) {
        case 1: TDZ_S16L_G(1); break;
        case 2: TDZ_S16L_G(2); break;
        case 3: TDZ_S16L_G(3); break;
        case 4: TDZ_S16L_G(4); break;
        default: TDZ_S16L_G(5); break;
        }
#undef TDZ_S16L_G
#undef TDZ_S16L
        return;
    }
    if (split16 == 2) {
        const uint4 *T4 = reinterpret_cast<const uint4 *>(Th);
        const bool kf = !DIAG && sK == 1 && nd % 2 == 0;
        if (kf ? launch_split16_lds<DIAG, true>(st, nd, ncol, Lin, sB, sL, sK, T4, ainv, x, y, dg, nbytes)
               : launch_split16_lds<DIAG, false>(st, nd, ncol, Lin, sB, sL, sK, T4, ainv, x, y, dg, nbytes))
            return;
    }
    const int nit = (nd + 15) / 16, G = nit < 3 ? nit : 3, ngr = (nit + G - 1) / G;
    const unsigned nblk = (unsigned)(((ncol + 15) / 16 * ngr + 3) / 4);
    const uint4 *T4 = reinterpret_cast<const uint4 *>(Th);
#define TDZ_S16(KB, GG)                                                                                             \
    do {                                                                                                        \
        if (!DIAG && sK == 1 && nd % 2 == 0)                                                                    \
            hipLaunchKernelGGL((axis_split16_kernel<KB, GG, false, 4, true>), dim3(nblk), dim3(256), 0, st, nd,  \
                               ncol, Lin, sB, sL, sK, nit, T4, ainv, x, y, dg, nbytes);                          \
        else                                                                                                    \
            hipLaunchKernelGGL((axis_split16_kernel<KB, GG, DIAG, 4, false>), dim3(nblk), dim3(256), 0, \
                               st, nd, ncol, Lin, sB, sL, sK, nit, T4, ainv, x, y, dg, nbytes);                  \
    } while (0)
#define TDZ_S16_G(KB)                 \
    do {                              \
        if (G == 1) TDZ_S16(KB, 1);   \
        else if (G == 2) TDZ_S16(KB, 2); \
        else TDZ_S16(KB, 3);          \
    } while (0)
    switch (axis_split16_blocks(nd)) {
    case 1: TDZ_S16_G(1); break;
    case 2: TDZ_S16_G(2); break;
    case 3: TDZ_S16_G(3); break;
    case 4: TDZ_S16_G(4); break;
    default: TDZ_S16_G(5); break;
    }
#undef TDZ_S16_G
#undef TDZ_S16
}

}  // namespace

bool launch_dense_axis_split16(hipStream_t st, int d, const int n[3], const unsigned short *Th, float ainv,
                               const fcomplex *x, fcomplex *y, int split16)
{
    if (!Th || !split16_fits(n) || !axis_split16_blocks(n[d])) return false;
    const long n01 = (long)n[0] * n[1];
    long L, sB, sK, sL;
    unsigned B;
    if (d == 0) { L = (long)n[1] * n[2]; sB = 0; sK = 1; sL = n[0]; B = 1; }
    else if (d == 1) { L = n[0]; sB = n01; sK = n[0]; sL = 1; B = (unsigned)n[2]; }
    else { L = n01; sB = 0; sK = n01; sL = 1; B = 1; }
    launch_split16_axis<false>(st, n[d], L * B, L, sB, sL, sK, Th, ainv, x, y, DiagF{}, (unsigned)(8 * n01 * n[2]),
                               split16);
    return true;
}

bool launch_dense_axis_diag_split16(hipStream_t st, const int n[3], const unsigned short *Th, float ainv,
                                    const float *lam[3], const unsigned char *wall[3], fcomplex alpha, fcomplex g,
                                    const fcomplex *x, fcomplex *y, int split16)
{
    if (!Th || !split16_fits(n) || !axis_split16_blocks(n[2])) return false;
    const long ncol = (long)n[0] * n[1];
    const DiagF dg{{lam[0], lam[1], lam[2]}, {wall[0], wall[1], wall[2]}, alpha, g, n[0]};
    launch_split16_axis<true>(st, n[2], ncol, ncol, 0, 1, ncol, Th, ainv, x, y, dg, (unsigned)(8 * ncol * n[2]),
                              split16);
    return true;
}

namespace {
// one launch of the fragment-ordered transform: KS from the axis length, G row tiles
// per wave (3, fewer on short axes)
template <bool DIAG>
void launch_wave_axis(hipStream_t st, int nd, long ncol, long Lin, long sB, long sL, long sK, const float *Tf,
                      const fcomplex *x, fcomplex *y, const DiagF &dg)
{
    const int nit = (nd + 15) / 16, G = nit < 3 ? nit : 3, ngr = (nit + G - 1) / G;
    const unsigned nblk = (unsigned)(((ncol + 15) / 16 * ngr + 3) / 4);
    const float4 *T4 = reinterpret_cast<const float4 *>(Tf);
#define TDZ_WAVE(KS, GG)                                                                                           \
    hipLaunchKernelGGL((axis_wave_mfma_kernel<KS, GG, DIAG>), dim3(nblk), dim3(256), 0, st, nd, ncol, Lin, sB, sL, sK, \
                       nit, T4, x, y, dg)
#define TDZ_WAVE_G(KS)              \
    do {                            \
        if (G == 1) TDZ_WAVE(KS, 1); \
        else if (G == 2) TDZ_WAVE(KS, 2); \
        else TDZ_WAVE(KS, 3);       \
    } while (0)
    switch (axis_frag_steps(nd)) {
    case 8: TDZ_WAVE_G(8); break;
    case 16: TDZ_WAVE_G(16); break;
    case 24: TDZ_WAVE_G(24); break;
    case 33: TDZ_WAVE_G(33); break;
    default: TDZ_WAVE_G(40); break;
    }
#undef TDZ_WAVE_G
#undef TDZ_WAVE
}
}  // namespace

template <typename R>
void launch_dense_axis(hipStream_t st, int d, const int n[3], const R *T, bool trans, const cpx<R> *x, cpx<R> *y,
                       const R *Tfrag)
{
    // batched GEMM view of the transform along axis d
    const long n01 = (long)n[0] * n[1];
    long L, sB, sK, sL;
    unsigned B;
    int kfast;
    if (d == 0) { L = (long)n[1] * n[2]; sB = 0; sK = 1; sL = n[0]; B = 1; kfast = 1; }
    else if (d == 1) { L = n[0]; sB = n01; sK = n[0]; sL = 1; B = (unsigned)n[2]; kfast = 0; }
    else { L = n01; sB = 0; sK = n01; sL = 1; B = 1; kfast = 0; }
    const dim3 grid((unsigned)((L + GT - 1) / GT), (unsigned)((n[d] + GT - 1) / GT), B);
    if constexpr (std::is_same<R, float>::value) {
        if (Tfrag && axis_frag_steps(n[d])) {   // columns across the batch: (c / Lin) sB + (c % Lin) sL
            launch_wave_axis<false>(st, n[d], L * B, L, sB, sL, sK, Tfrag, x, y, DiagF{});
            return;
        }
        hipLaunchKernelGGL(axis_gemm_mfma_kernel, grid, dim3(256), 0, st, n[d], L, sB, sK, sL, kfast, T,
                           trans ? 1 : 0, x, y);
    } else {
        const int KS = axis_f64_steps(n[d]);
        if (Tfrag && KS) {   // FP64 MFMA, a wave per (column tile, row tile)
            const int nit = (n[d] + 15) / 16;
            const long ncol = L * B;
            const unsigned nblk = (unsigned)(((ncol + 15) / 16 * nit + 3) / 4);
            const double2 *T2 = reinterpret_cast<const double2 *>(Tfrag);
#define TDZ_F64(K)                                                                                                 \
    case K:                                                                                                        \
        hipLaunchKernelGGL(axis_wave_f64_kernel<K>, dim3(nblk), dim3(256), 0, st, n[d], ncol, L, sB, sL, sK, nit, T2, \
                           x, y);                                                                                  \
        return;
            switch (KS) {
                TDZ_F64(2) TDZ_F64(4) TDZ_F64(6) TDZ_F64(8) TDZ_F64(10) TDZ_F64(12) TDZ_F64(14) TDZ_F64(16)
                TDZ_F64(18) TDZ_F64(20) TDZ_F64(22) TDZ_F64(24) TDZ_F64(26) TDZ_F64(28) TDZ_F64(30) TDZ_F64(32)
                TDZ_F64(34) TDZ_F64(36) TDZ_F64(38) TDZ_F64(40)
            default: break;
            }
#undef TDZ_F64
        }
        hipLaunchKernelGGL(axis_gemm_kernel<R>, grid, dim3(256), 0, st, n[d], L, sB, sK, sL, kfast, T,
                           trans ? 1 : 0, x, y);
    }
}

bool launch_dense_axis_diag(hipStream_t st, const int n[3], const float *Tfrag, const float *lam[3],
                            const unsigned char *wall[3], fcomplex alpha, fcomplex g, const fcomplex *x, fcomplex *y)
{
    if (!Tfrag || !axis_frag_steps(n[2])) return false;
    const long ncol = (long)n[0] * n[1];
    const DiagF dg{{lam[0], lam[1], lam[2]}, {wall[0], wall[1], wall[2]}, alpha, g, n[0]};
    launch_wave_axis<true>(st, n[2], ncol, ncol, 0, 1, ncol, Tfrag, x, y, dg);
    return true;
}

template <typename R>
void launch_dense_axis_simple(hipStream_t st, int d, const int n[3], const R *T, bool trans, const cpx<R> *x,
                              cpx<R> *y)
{
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(dense_axis_kernel<R>, dim3(grid_for(N, 256)), dim3(256), 0, st, d, n[0], n[1], n[2], T,
                       trans ? 1 : 0, x, y);
}

template <typename R>
void launch_diag_scale(hipStream_t st, const int n[3], const R *lam[3], const unsigned char *wall[3], cpx<R> alpha,
                       cpx<R> g, const cpx<R> *x, cpx<R> *y)
{
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(diag_scale_kernel<R>, dim3(grid_for(N, 256)), dim3(256), 0, st, n[0], n[1], N, lam[0], lam[1],
                       lam[2], wall[0], wall[1], wall[2], alpha, g, x, y);
}

template <typename R>
void launch_csr_axis(hipStream_t st, int d, const int nin[3], int m, const int *rowptr, const int *col,
                     const R *val, const cpx<R> *in, cpx<R> *out, bool accumulate)
{
    const long N = (long)(d == 0 ? m : nin[0]) * (d == 1 ? m : nin[1]) * (d == 2 ? m : nin[2]);
    hipLaunchKernelGGL(csr_axis_kernel<R>, dim3(grid_for(N, 256)), dim3(256), 0, st, d, nin[0], nin[1], nin[2], m,
                       rowptr, col, val, in, out, accumulate ? 1 : 0);
}

// ------------------------------------------------------------ device-side FGMRES(k)
// The small Hessenberg least-squares problem of a fixed-length FGMRES cycle on the
// device (one thread), so that a whole time step can be captured in a HIP graph
// without host round trips; same rotations as the host solver.
namespace {
__device__ __forceinline__ dcomplex cmul(dcomplex a, dcomplex b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ dcomplex cconj(dcomplex a) { return {a.re, -a.im}; }
__device__ __forceinline__ dcomplex cadd(dcomplex a, dcomplex b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ dcomplex csub(dcomplex a, dcomplex b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ double cabs2(dcomplex a) { return a.re * a.re + a.im * a.im; }

__global__ void gm_start_kernel(const double *beta2, const double *bnorm2, GmState *g)
{
    g->gv[0] = {sqrt(beta2[0]), 0.0};
    for (int i = 1; i <= kGmMax; ++i) g->gv[i] = {0.0, 0.0};
    g->bnorm = sqrt(bnorm2[0]);
    g->done = 0;
}

__global__ void gm_arnoldi_kernel(int j, int m, const dcomplex *h1, const dcomplex *h2, const double *wn2, GmState *g)
{
    dcomplex *H = g->H;   // [(m + 1) x m] row-major, column j
    for (int i = 0; i <= j; ++i) H[i * m + j] = cadd(h1[i], h2[i]);
    const double hn = sqrt(wn2[0]);
    H[(j + 1) * m + j] = {hn, 0.0};
    for (int i = 0; i < j; ++i) {
        const dcomplex a = H[i * m + j], b = H[(i + 1) * m + j];
        const double c = g->cs[i];
        const dcomplex sn = g->sn[i];
        H[i * m + j] = cadd({c * a.re, c * a.im}, cmul(sn, b));
        const dcomplex t = cmul(cconj(sn), a);
        H[(i + 1) * m + j] = {c * b.re - t.re, c * b.im - t.im};
    }
    const dcomplex h1j = H[j * m + j], h2j = H[(j + 1) * m + j];
    const double rho = sqrt(cabs2(h1j) + cabs2(h2j));
    const double a1 = sqrt(cabs2(h1j));
    if (a1 == 0.0) {
        g->cs[j] = 0.0;
        g->sn[j] = {1.0, 0.0};
    } else {
        g->cs[j] = a1 / rho;
        const dcomplex ph = {h1j.re / a1, h1j.im / a1};
        const dcomplex t = cmul(ph, cconj(h2j));
        g->sn[j] = {t.re / rho, t.im / rho};
    }
    const double c = g->cs[j];
    const dcomplex sn = g->sn[j];
    H[j * m + j] = cadd({c * h1j.re, c * h1j.im}, cmul(sn, h2j));
    H[(j + 1) * m + j] = {0.0, 0.0};
    const dcomplex gj = g->gv[j];
    const dcomplex t = cmul(cconj(sn), gj);
    g->gv[j + 1] = {-t.re, -t.im};
    g->gv[j] = {c * gj.re, c * gj.im};
}

// y = R^-1 g over the first k columns; res[*step] = |g_k| / ||b||
__global__ void gm_solve_kernel(int k, int m, GmState *g, dcomplex *y, double *res, const int *step)
{
    const dcomplex *H = g->H;
    for (int i = k - 1; i >= 0; --i) {
        dcomplex s = g->gv[i];
        for (int l = i + 1; l < k; ++l) s = csub(s, cmul(H[i * m + l], y[l]));
        const dcomplex d = H[i * m + i];
        const double dd = cabs2(d);
        y[i] = dd > 0.0 ? dcomplex{(s.re * d.re + s.im * d.im) / dd, (s.im * d.re - s.re * d.im) / dd}
                        : dcomplex{0.0, 0.0};
    }
    res[*step] = g->bnorm > 0.0 ? sqrt(cabs2(g->gv[k])) / g->bnorm : 0.0;
}

__global__ void step_advance_kernel(int *step) { ++step[0]; }

// B[i] = (M ? M[i] : 0) + g (T0[i] + sE[3 step + d] X[i]) (band of one axis for this step's field)
__global__ void field_band_kernel(long n, const double *__restrict__ M, const dcomplex *__restrict__ T0,
                                  const double *__restrict__ X, dcomplex g, const double *__restrict__ sE,
                                  const int *__restrict__ step, int d, dcomplex *__restrict__ B)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double e = sE[3 * (long)step[0] + d];
    const dcomplex t = {T0[i].re + e * X[i], T0[i].im};
    const dcomplex gt = cmul(g, t);
    B[i] = {(M ? M[i] : 0.0) + gt.re, gt.im};
}
}  // namespace

void launch_gm_start(hipStream_t st, const double *beta2, const double *bnorm2, GmState *g)
{
    hipLaunchKernelGGL(gm_start_kernel, dim3(1), dim3(1), 0, st, beta2, bnorm2, g);
}
void launch_gm_arnoldi(hipStream_t st, int j, int m, const dcomplex *h1, const dcomplex *h2, const double *wn2,
                       GmState *g)
{
    hipLaunchKernelGGL(gm_arnoldi_kernel, dim3(1), dim3(1), 0, st, j, m, h1, h2, wn2, g);
}
void launch_gm_solve(hipStream_t st, int k, int m, GmState *g, dcomplex *y, double *res, const int *step)
{
    hipLaunchKernelGGL(gm_solve_kernel, dim3(1), dim3(1), 0, st, k, m, g, y, res, step);
}
void launch_step_advance(hipStream_t st, int *step)
{
    hipLaunchKernelGGL(step_advance_kernel, dim3(1), dim3(1), 0, st, step);
}
void launch_field_band(hipStream_t st, long n, const double *M, const dcomplex *T0, const double *X, dcomplex g,
                       const double *sE, const int *step, int d, dcomplex *B)
{
    hipLaunchKernelGGL(field_band_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, M, T0, X, g, sE,
                       step, d, B);
}

// ------------------------------------------------------------ z-line solves
namespace {
template <typename R>
__device__ __forceinline__ cpx<R> cmulR(cpx<R> a, cpx<R> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <typename R>
__device__ __forceinline__ cpx<R> cdivR(cpx<R> a, cpx<R> b)
{
    const R d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

template <typename R>
__global__ void zline_solve_kernel(int n0, int n1, int n2, int p, const R *__restrict__ lx, const R *__restrict__ ly,
                                   const cpx<R> *__restrict__ Bz, const R *__restrict__ Mw, cpx<R> g,
                                   const cpx<R> *__restrict__ rhs, cpx<R> *__restrict__ x, cpx<R> *__restrict__ work)
{
    const long nl = (long)n0 * n1;
    const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nl) return;
    const int i = (int)(line % n0), j = (int)(line / n0);
    const R lam = lx[i] + ly[j];
    const cpx<R> cg = {g.re * lam, g.im * lam};
    const int bw = 2 * p + 1;
    auto W = [&](int k, int m) -> cpx<R> & { return work[((long)k * bw + m) * nl + line]; };
    for (int k = 0; k < n2; ++k) {
        for (int m = 0; m < bw; ++m) {
            const int col = k - p + m;
            cpx<R> v = {0, 0};
            if (col >= 0 && col < n2) {
                const cpx<R> b = Bz[(long)k * bw + m];
                const R w = Mw[(long)k * bw + m];
                v = {b.re + cg.re * w, b.im + cg.im * w};
            }
            W(k, m) = v;
        }
        x[line + nl * k] = rhs[line + nl * k];
    }
    for (int k = 0; k < n2; ++k) {   // elimination below the diagonal (no fill-in without pivoting)
        const cpx<R> piv = W(k, p);
        const cpx<R> xk = x[line + nl * k];
        const int rmax = k + p < n2 - 1 ? k + p : n2 - 1;
        for (int r = k + 1; r <= rmax; ++r) {
            const cpx<R> f = cdivR(W(r, k - r + p), piv);
            for (int m2 = 1; m2 <= p && k + m2 < n2; ++m2) {
                const cpx<R> t = cmulR(f, W(k, p + m2));
                cpx<R> &d = W(r, k + m2 - r + p);
                d = {d.re - t.re, d.im - t.im};
            }
            const cpx<R> t = cmulR(f, xk);
            cpx<R> &xr = x[line + nl * r];
            xr = {xr.re - t.re, xr.im - t.im};
        }
    }
    for (int k = n2 - 1; k >= 0; --k) {   // back substitution
        cpx<R> s = x[line + nl * k];
        for (int m2 = 1; m2 <= p && k + m2 < n2; ++m2) {
            const cpx<R> t = cmulR(W(k, p + m2), x[line + nl * (k + m2)]);
            s = {s.re - t.re, s.im - t.im};
        }
        x[line + nl * k] = cdivR(s, W(k, p));
    }
}

// The same banded elimination with the active rows in registers: rows k .. k+P of the
// matrix, columns k .. k+2P, slide down the line; row k+P+1 enters from the band tables
// (shared by all lines, cached), the finished row k of U and the eliminated x_k go to
// `work` and `x` once, and the back substitution reads them once. The operations and
// their order are those of zline_solve_kernel; only the storage moves out of the
// dependency chain (global-memory round trips per update made the line solves
// latency-bound: 1.5 ms at 68 x 68 x 260).
template <typename R, int P>
__global__ void __launch_bounds__(64) zline_reg_kernel(int n0, int n1, int n2, const R *__restrict__ lx,
                                                      const R *__restrict__ ly,
                                 const cpx<R> *__restrict__ Bz, const R *__restrict__ Mw, cpx<R> g,
                                 const cpx<R> *__restrict__ rhs, cpx<R> *__restrict__ x, cpx<R> *__restrict__ work)
{
    constexpr int BW = 2 * P + 1;
    const long nl = (long)n0 * n1;
    const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nl) return;
    const int i = (int)(line % n0), j = (int)(line / n0);
    const R lam = lx[i] + ly[j];
    const cpx<R> cg = {g.re * lam, g.im * lam};
    auto entry = [&](int r, int col) -> cpx<R> {   // A(r, col), zero outside the matrix and the band
        if (r >= n2 || col < 0 || col >= n2 || col - r > P || r - col > P) return {0, 0};
        const cpx<R> b = Bz[(long)r * BW + (col - r + P)];
        const R w = Mw[(long)r * BW + (col - r + P)];
        return {b.re + cg.re * w, b.im + cg.im * w};
    };
    cpx<R> W[P + 1][BW];   // W[rr][cc] = A(k + rr, k + cc)
    cpx<R> xw[P + 1];      // x at rows k .. k+P
#pragma unroll
    for (int rr = 0; rr <= P; ++rr) {
#pragma unroll
        for (int cc = 0; cc < BW; ++cc) W[rr][cc] = entry(rr, cc);
        xw[rr] = rr < n2 ? rhs[line + nl * rr] : cpx<R>{0, 0};
    }
    for (int k = 0; k < n2; ++k) {
        const cpx<R> piv = W[0][0];
#pragma unroll
        for (int rr = 1; rr <= P; ++rr) {
            if (k + rr >= n2) break;
            const cpx<R> f = cdivR(W[rr][0], piv);
#pragma unroll
            for (int m2 = 1; m2 <= P; ++m2) {
                if (k + m2 >= n2) break;
                const cpx<R> t = cmulR(f, W[0][m2]);
                W[rr][m2] = {W[rr][m2].re - t.re, W[rr][m2].im - t.im};
            }
            const cpx<R> t = cmulR(f, xw[0]);
            xw[rr] = {xw[rr].re - t.re, xw[rr].im - t.im};
        }
        // row k of U (diagonal and P super-diagonals) and the eliminated x_k
#pragma unroll
        for (int m = 0; m <= P; ++m) work[((long)k * (P + 1) + m) * nl + line] = W[0][m];
        x[line + nl * k] = xw[0];
        // slide: rows k+1 .. k+P move up one row and left one column, row k+P+1 enters
#pragma unroll
        for (int rr = 0; rr < P; ++rr) {
#pragma unroll
            for (int cc = 0; cc < BW - 1; ++cc) W[rr][cc] = W[rr + 1][cc + 1];
            W[rr][BW - 1] = {0, 0};
            xw[rr] = xw[rr + 1];
        }
        const int rn = k + P + 1;
#pragma unroll
        for (int cc = 0; cc < BW; ++cc) W[P][cc] = entry(rn, k + 1 + cc);
        xw[P] = rn < n2 ? rhs[line + nl * rn] : cpx<R>{0, 0};
    }
    cpx<R> xs[P];   // x at rows k+1 .. k+P of the back substitution
#pragma unroll
    for (int m = 0; m < P; ++m) xs[m] = {0, 0};
    for (int k = n2 - 1; k >= 0; --k) {
        cpx<R> s = x[line + nl * k];
#pragma unroll
        for (int m2 = 1; m2 <= P; ++m2) {
            if (k + m2 >= n2) break;
            const cpx<R> t = cmulR(work[((long)k * (P + 1) + m2) * nl + line], xs[m2 - 1]);
            s = {s.re - t.re, s.im - t.im};
        }
        const cpx<R> xk = cdivR(s, work[((long)k * (P + 1)) * nl + line]);
        x[line + nl * k] = xk;
#pragma unroll
        for (int m = P - 1; m > 0; --m) xs[m] = xs[m - 1];
        xs[0] = xk;
    }
}
}  // namespace

template <typename R>
void launch_zline_solve(hipStream_t st, const int n[3], int p, const R *lx, const R *ly, const cpx<R> *Bz,
                        const R *Mw, cpx<R> g, const cpx<R> *rhs, cpx<R> *x, cpx<R> *work)
{
    const long nl = (long)n[0] * n[1];
    const dim3 grid((unsigned)((nl + 63) / 64));
    switch (p) {   // work holds n2 (p+1) entries per line here (of the n2 (2p+1) allocated)
    case 1: hipLaunchKernelGGL((zline_reg_kernel<R, 1>), grid, dim3(64), 0, st, n[0], n[1], n[2], lx, ly, Bz, Mw, g, rhs, x, work); return;
    case 2: hipLaunchKernelGGL((zline_reg_kernel<R, 2>), grid, dim3(64), 0, st, n[0], n[1], n[2], lx, ly, Bz, Mw, g, rhs, x, work); return;
    case 3: hipLaunchKernelGGL((zline_reg_kernel<R, 3>), grid, dim3(64), 0, st, n[0], n[1], n[2], lx, ly, Bz, Mw, g, rhs, x, work); return;
    case 4: hipLaunchKernelGGL((zline_reg_kernel<R, 4>), grid, dim3(64), 0, st, n[0], n[1], n[2], lx, ly, Bz, Mw, g, rhs, x, work); return;
    case 5: hipLaunchKernelGGL((zline_reg_kernel<R, 5>), grid, dim3(64), 0, st, n[0], n[1], n[2], lx, ly, Bz, Mw, g, rhs, x, work); return;
    default: break;
    }
    hipLaunchKernelGGL(zline_solve_kernel<R>, grid, dim3(64), 0, st, n[0], n[1], n[2], p, lx, ly, Bz, Mw, g, rhs, x,
                       work);
}

template <typename R>
void launch_jacobi_diag(hipStream_t st, const int n[3], const int p[3], const R *M[3], const cpx<R> *B[3],
                        const R *Vc, const R *Vl, cpx<R> g, const unsigned char *wall[3], cpx<R> *dinv)
{
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(jacobi_diag_kernel<R>, dim3((unsigned)std::min<long>((N + 255) / 256, 1L << 20)), dim3(256), 0,
                       st, n[0], n[1], n[2], p[0], p[1], p[2], M[0], M[1], M[2], B[0], B[1], B[2], Vc, Vl, g.re, g.im,
                       wall[0], wall[1], wall[2], dinv);
}

template <typename R>
void launch_cheb_update(hipStream_t st, long N, R a, R b, const cpx<R> *dinv, const cpx<R> *res, cpx<R> *d,
                        cpx<R> *e)
{
    hipLaunchKernelGGL(cheb_update_kernel<R>, dim3((unsigned)std::min<long>((N + 255) / 256, 1L << 20)), dim3(256), 0,
                       st, N, a, b, dinv, res, d, e);
}

#define TDZ_INSTANTIATE(R)                                                                                          \
    template void launch_jacobi_diag<R>(hipStream_t, const int[3], const int[3], const R *[3], const cpx<R> *[3],   \
                                        const R *, const R *, cpx<R>, const unsigned char *[3], cpx<R> *);         \
    template void launch_cheb_update<R>(hipStream_t, long, R, R, const cpx<R> *, const cpx<R> *, cpx<R> *,         \
                                        cpx<R> *);                                                                 \
    template void launch_pot_apply_generic<R>(hipStream_t, const AxisDevT<R>[3], PotParams, R, R, const cpx<R> *,  \
                                              cpx<R> *);                                                           \
    template bool launch_pot_apply_3d<R>(hipStream_t, const AxisDevT<R>[3], PotParams, R, R, const cpx<R> *,       \
                                         cpx<R> *, int, cpx<R> *);                                                 \
    template bool launch_pot_apply_2d<R>(hipStream_t, const AxisDevT<R>[3], PotParams, R, R, const cpx<R> *,       \
                                         cpx<R> *, int);                                                           \
    template void launch_sep_apply<R>(hipStream_t, const SepBandsT<R> &, const cpx<R> *, cpx<R> *, cpx<R> *[4]);   \
    template void launch_wall_fix<R>(hipStream_t, const int[3], const unsigned char *[3], const cpx<R> *,          \
                                     cpx<R> *);                                                                    \
    template void launch_axpby_t<R>(hipStream_t, cpx<R>, const cpx<R> *, cpx<R>, cpx<R> *, long);                  \
    template void launch_dense_axis<R>(hipStream_t, int, const int[3], const R *, bool, const cpx<R> *, cpx<R> *,   \
                                        const R *);                                                               \
    template void launch_dense_axis_simple<R>(hipStream_t, int, const int[3], const R *, bool, const cpx<R> *,     \
                                              cpx<R> *);                                                           \
    template void launch_diag_scale<R>(hipStream_t, const int[3], const R *[3], const unsigned char *[3], cpx<R>,  \
                                       cpx<R>, const cpx<R> *, cpx<R> *);                                          \
    template void launch_csr_axis<R>(hipStream_t, int, const int[3], int, const int *, const int *, const R *,     \
                                     const cpx<R> *, cpx<R> *, bool);                                              \
    template void launch_zline_solve<R>(hipStream_t, const int[3], int, const R *, const R *, const cpx<R> *,      \
                                        const R *, cpx<R>, const cpx<R> *, cpx<R> *, cpx<R> *);

TDZ_INSTANTIATE(double)
TDZ_INSTANTIATE(float)

}  // namespace tdz
