// Exact Galerkin coarse potential for the multigrid levels (gfx950).
//
// The reference coarsens the whole propagator operator by PETSc's PtAP
// (PCMGSetGalerkin(pc, PC_MG_GALERKIN_BOTH), src/multigrid.cpp:277). The separable
// terms coarsen per axis (band products); the potential does not, so it is assembled
// here once per run as a variable-coefficient stencil of the coarse level:
//
//   V_l(i, i+d) = sum over the finest level's Gauss points q of
//                 w_q s V(x_q) prod_a N^l_{i_a}(x_qa) N^l_{i_a+d_a}(x_qa),
//
// i.e. P^T V_0 P with P the composite knot-insertion prolongation (the coarse
// B-splines are exact combinations of the fine ones, so their products at the fine
// points reproduce P^T V_0 P to rounding). Per axis a "pair table" holds, for every
// (function i, offset d) pair, the products N_i N_{i+d} at the contiguous run of fine
// points where both are nonzero. The 3D sum is three contractions, one axis at a time
// (x with the potential evaluated on the fly, then y, then z), over slabs of fine z
// points so the intermediates stay bounded.
//
// The stencil is stored offset-major: entry (i, d) at
//   ((dz nz + iz) Dy + dy) ny + iy) Dx nx + dx nx + ix,   D_a = 2p+1 (1 on unused axes),
// so for a fixed offset consecutive ix are consecutive in memory (coalesced apply).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "tdz_internal.hpp"

namespace tdz {

namespace {

__device__ __forceinline__ double gal_pot(const PotParams &pp, double r2)
{
    if (pp.kind == 1) return 0.5 * pp.a0 * r2;
    if (pp.kind == 2) return -pp.a0 / sqrt(r2 + pp.a1);
    return 0.0;
}

// T1[(zl ny_q + qy) Px + px] = sum_k val[off[px] + k] W(q0[px] + k, qy, z0 + zl),
// W = wx wy wz s V(x, y, z) at the finest level's Gauss points.
__global__ void __launch_bounds__(256) gal_pass_x_kernel(GalPairs tx, const double *__restrict__ X,
                                                         const double *__restrict__ WX, const double *__restrict__ Y,
                                                         const double *__restrict__ WY, const double *__restrict__ Z,
                                                         const double *__restrict__ WZ, int nqy, int z0, int nz,
                                                         PotParams pp, double *__restrict__ T1)
{
    const long total = (long)nz * nqy * tx.npairs;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int px = (int)(t % tx.npairs);
        const long row = t / tx.npairs;
        const int qy = (int)(row % nqy), zl = (int)(row / nqy);
        const int qz = z0 + zl;
        const double yz2 = Y[qy] * Y[qy] + Z[qz] * Z[qz], wyz = WY[qy] * WZ[qz] * pp.s;
        const int q0 = tx.q0[px], len = tx.len[px];
        const double *v = tx.val + tx.off[px];
        double acc = 0.0;
        for (int k = 0; k < len; ++k) {
            const double xq = X[q0 + k];
            acc += v[k] * (WX[q0 + k] * gal_pot(pp, xq * xq + yz2));
        }
        T1[t] = acc * wyz;
    }
}

// out[(a P + p) B + b] (+)= sum_k val[off[p] + k] in[(a Qn + q) B + b], q = q0[p] + k - qbase,
// over the q of [0, Qn): one contraction of the middle index of in[A][Qn][B].
__global__ void __launch_bounds__(256) gal_contract_kernel(GalPairs t, int A, int Qn, int qbase, long B,
                                                           const double *__restrict__ in, double *__restrict__ out,
                                                           int accumulate)
{
    const long total = (long)A * t.npairs * B;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long b = i % B;
        const long ap = i / B;
        const int p = (int)(ap % t.npairs), a = (int)(ap / t.npairs);
        const int lo = max(t.q0[p] - qbase, 0), hi = min(t.q0[p] + t.len[p] - qbase, Qn);
        const double *v = t.val + t.off[p] + (qbase + lo - t.q0[p]);
        const double *src = in + ((long)a * Qn + lo) * B + b;
        double acc = 0.0;
        for (int q = lo; q < hi; ++q, ++v, src += B) acc += *v * *src;
        out[i] = accumulate ? out[i] + acc : acc;
    }
}

__global__ void gal_demote_kernel(const double *__restrict__ s, float *__restrict__ f, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        f[i] = (float)s[i];
}

// y[I] += g sum_d S[I, d] x[I + d], one thread per row I (x fastest)
template <typename R>
__global__ void __launch_bounds__(256) stencil_apply_kernel(int nx, int ny, int nz, int px, int py, int pz,
                                                            const R *__restrict__ S, R gr, R gi,
                                                            const cpx<R> *__restrict__ x, cpx<R> *__restrict__ y)
{
    const long N = (long)nx * ny * nz;
    const int Dx = 2 * px + 1, Dy = 2 * py + 1;
    for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < N; I += (long)gridDim.x * blockDim.x) {
        const int ix = (int)(I % nx);
        const int iy = (int)((I / nx) % ny);
        const int iz = (int)(I / ((long)nx * ny));
        R ar = 0, ai = 0;
        for (int dz = max(-pz, -iz); dz <= min(pz, nz - 1 - iz); ++dz)
            for (int dy = max(-py, -iy); dy <= min(py, ny - 1 - iy); ++dy) {
                const long srow = ((((long)(dz + pz) * nz + iz) * Dy + (dy + py)) * ny + iy) * Dx * nx;
                const long xrow = ((long)(iz + dz) * ny + (iy + dy)) * nx + ix;
                for (int dx = max(-px, -ix); dx <= min(px, nx - 1 - ix); ++dx) {
                    const R s = S[srow + (long)(dx + px) * nx + ix];
                    const cpx<R> v = x[xrow + dx];
                    ar += s * v.re;
                    ai += s * v.im;
                }
            }
        cpx<R> o = y[I];
        o.re += gr * ar - gi * ai;
        o.im += gr * ai + gi * ar;
        y[I] = o;
    }
}

// The same for a 3D stencil of half-width P on every axis. A block of 64 x D threads takes
// 64 consecutive rows; thread (r, dz) sums the (2P+1)^2 terms of z offset dz of row r
// (fully unrolled: every load of the slab is in flight at once), and the D partial sums
// of a row are added in LDS in a fixed order. Neighbour indices are clamped into the
// grid: a clamped neighbour carries a zero stencil entry (no pair of functions at that
// offset), so the loops have no branches.
template <typename R, int P>
__global__ void __launch_bounds__(64 * (2 * P + 1)) stencil3_kernel(int nx, int ny, int nz, const R *__restrict__ S,
                                                                    R gr, R gi, const cpx<R> *__restrict__ x,
                                                                    cpx<R> *__restrict__ y)
{
    constexpr int D = 2 * P + 1;
    __shared__ R part[2][D][64];
    const long N = (long)nx * ny * nz;
    const long sX = nx, sY = (long)ny * D * nx, sZ = (long)nz * D * sY;   // strides of dx', dy', dz'
    const int r = threadIdx.x, dz = threadIdx.y;
    const long I = blockIdx.x * 64L + r;
    R ar = 0, ai = 0;
    if (I < N) {
        const int ix = (int)(I % nx);
        const int iy = (int)((I / nx) % ny);
        const int iz = (int)(I / ((long)nx * ny));
        const R *Sz = S + dz * sZ + (long)iz * D * sY + (long)iy * D * nx + ix;
        const cpx<R> *xz = x + (long)min(max(iz + dz - P, 0), nz - 1) * ny * nx;
#pragma unroll
        for (int dy = 0; dy < D; ++dy) {
            const cpx<R> *xr = xz + (long)min(max(iy + dy - P, 0), ny - 1) * nx;
#pragma unroll
            for (int dx = 0; dx < D; ++dx) {
                const R sv = Sz[dy * sY + dx * sX];
                const cpx<R> v = xr[min(max(ix + dx - P, 0), nx - 1)];
                ar += sv * v.re;
                ai += sv * v.im;
            }
        }
    }
    part[0][dz][r] = ar;
    part[1][dz][r] = ai;
    __syncthreads();
    if (dz == 0 && I < N) {
        R sr = 0, si = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            sr += part[0][k][r];
            si += part[1][k][r];
        }
        cpx<R> o = y[I];
        o.re += gr * sr - gi * si;
        o.im += gr * si + gi * sr;
        y[I] = o;
    }
}

inline int grid_for(long n, int block = 256)
{
    const long b = (n + block - 1) / block;
    return (int)std::min<long>(std::max<long>(b, 1), 1L << 20);
}

}  // namespace

void galerkin_potential(hipStream_t st, const GalerkinPlan &g, double *S, double *work, long work_len)
{
    const int Px = g.pairs[0].npairs, Py = g.pairs[1].npairs, Pz = g.pairs[2].npairs;
    const int nqx = g.nq[0], nqy = g.nq[1], nqz = g.nq[2];
    (void)nqx;
    // z slab: T1 [cz][nqy][Px] and T2 [cz][Py][Px] in the work buffer
    const long per_z = (long)nqy * Px + (long)Py * Px;
    const int cz = (int)std::min<long>(nqz, std::max<long>(1, work_len / per_z));
    if (work_len < per_z) throw std::runtime_error("galerkin_potential: work buffer too small");
    double *T1 = work, *T2 = work + (long)cz * nqy * Px;
    for (int z0 = 0; z0 < nqz; z0 += cz) {
        const int nz = std::min(cz, nqz - z0);
        const long n1 = (long)nz * nqy * Px;
        gal_pass_x_kernel<<<grid_for(n1), 256, 0, st>>>(g.pairs[0], g.X[0], g.W[0], g.X[1], g.W[1], g.X[2], g.W[2],
                                                       nqy, z0, nz, g.pot, T1);
        const long n2 = (long)nz * Py * Px;
        gal_contract_kernel<<<grid_for(n2), 256, 0, st>>>(g.pairs[1], nz, nqy, 0, Px, T1, T2, 0);
        const long n3 = (long)Pz * Py * Px;
        gal_contract_kernel<<<grid_for(n3), 256, 0, st>>>(g.pairs[2], 1, nz, z0, (long)Py * Px, T2, S, z0 > 0);
    }
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("galerkin_potential: launch failed");
}

void launch_stencil_demote(hipStream_t st, const double *S, float *Sf, long n)
{
    gal_demote_kernel<<<grid_for(n), 256, 0, st>>>(S, Sf, n);
}

template <typename R>
void launch_stencil_apply(hipStream_t st, const int n[3], const int p[3], const R *S, R gr, R gi, const cpx<R> *x,
                          cpx<R> *y)
{
    const long N = (long)n[0] * n[1] * n[2];
    const bool cube = p[0] == p[1] && p[1] == p[2] && n[0] > 1 && n[1] > 1 && n[2] > 1;
    const unsigned nb = (unsigned)((N + 63) / 64);
    if (cube) {
        switch (p[0]) {
        case 2: stencil3_kernel<R, 2><<<nb, dim3(64, 5), 0, st>>>(n[0], n[1], n[2], S, gr, gi, x, y); return;
        case 3: stencil3_kernel<R, 3><<<nb, dim3(64, 7), 0, st>>>(n[0], n[1], n[2], S, gr, gi, x, y); return;
        case 4: stencil3_kernel<R, 4><<<nb, dim3(64, 9), 0, st>>>(n[0], n[1], n[2], S, gr, gi, x, y); return;
        case 5: stencil3_kernel<R, 5><<<nb, dim3(64, 11), 0, st>>>(n[0], n[1], n[2], S, gr, gi, x, y); return;
        default: break;
        }
    }
    stencil_apply_kernel<R><<<grid_for(N), 256, 0, st>>>(n[0], n[1], n[2], p[0], p[1], p[2], S, gr, gi, x, y);
}

template void launch_stencil_apply<double>(hipStream_t, const int *, const int *, const double *, double, double,
                                           const dcomplex *, dcomplex *);
template void launch_stencil_apply<float>(hipStream_t, const int *, const int *, const float *, float, float,
                                          const fcomplex *, fcomplex *);

}  // namespace tdz
