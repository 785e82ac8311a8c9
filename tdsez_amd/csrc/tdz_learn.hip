// Kernels of the learned potential-product network (gfx950).
//
// The network maps CI real input fields u_c on the coefficient grid (one component
// of CI complex64 vectors at a time, u_c at u + c N) to
//   out(v) = sum_b P[v][b] sum_t sum_a W2[t][a][b] m[v+t][a],
//   m[v][a] = P[v][a] sum_t sum_c W1[t][c][a] u_c(v+t),
// 3x3x3 taps t, C = 16 hidden channels, zero padding, P fixed positional fields.
//
//   cnn_l1 / cnn_l2     FP32 VALU forward (training and reference)
//   cnn_dm, cnn_dw*     FP32 VALU backward (gradients of the normalised L2 loss)
//   cnn_infer_*         BF16 MFMA forward with FP32 accumulation (inference in
//                       the preconditioner), v_mfma_f32_16x16x16_bf16
#include <hip/hip_runtime.h>

#include "tdz_internal.hpp"

namespace tdz {

namespace {

constexpr int C = kCnnChannels;
constexpr int CI = kCnnInputs;
constexpr int NTAP = 27;
constexpr int K1 = ((NTAP * CI + 15) / 16) * 16;   // layer-1 reduction length padded for MFMA

__device__ __forceinline__ int tap_dx(int t) { return t % 3 - 1; }
__device__ __forceinline__ int tap_dy(int t) { return (t / 3) % 3 - 1; }
__device__ __forceinline__ int tap_dz(int t) { return t / 9 - 1; }

__device__ __forceinline__ float comp_of(const fcomplex v, int comp) { return comp ? v.im : v.re; }

// neighbour index of voxel (i, j, k) by tap t, or -1 outside the grid
__device__ __forceinline__ long nb(int i, int j, int k, int t, int n0, int n1, int n2)
{
    const int a = i + tap_dx(t), b = j + tap_dy(t), c = k + tap_dz(t);
    if (a < 0 || a >= n0 || b < 0 || b >= n1 || c < 0 || c >= n2) return -1;
    return a + (long)n0 * (b + (long)n1 * c);
}

// ---------------------------------------------------------------- FP32 forward
// m[(comp N + v) C + a] = P[v][a] sum_t sum_c W1[t][c][a] u_c,comp(v + t)
__global__ void __launch_bounds__(256) cnn_l1_kernel(int n0, int n1, int n2, const fcomplex *__restrict__ u,
                                                     const float *__restrict__ W1, const float *__restrict__ P,
                                                     float *__restrict__ m)
{
    __shared__ float w[NTAP * CI * C];
    for (int i = threadIdx.x; i < NTAP * CI * C; i += blockDim.x) w[i] = W1[i];
    __syncthreads();
    const long N = (long)n0 * n1 * n2;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= 2 * N) return;
    const int comp = (int)(I / N);
    const long v = I % N;
    const int i = (int)(v % n0), j = (int)((v / n0) % n1), k = (int)(v / ((long)n0 * n1));
    float h[C];
#pragma unroll
    for (int a = 0; a < C; ++a) h[a] = 0.f;
    for (int t = 0; t < NTAP; ++t) {
        const long q = nb(i, j, k, t, n0, n1, n2);
        if (q < 0) continue;
#pragma unroll
        for (int c = 0; c < CI; ++c) {
            const float x = comp_of(u[(long)c * N + q], comp);
#pragma unroll
            for (int a = 0; a < C; ++a) h[a] += w[(t * CI + c) * C + a] * x;
        }
    }
#pragma unroll
    for (int a = 0; a < C; ++a) m[I * C + a] = P[v * C + a] * h[a];
}

// out[comp N + v] = sum_b P[v][b] sum_t sum_a W2[t][a][b] m[comp][v + t][a]
__global__ void __launch_bounds__(256) cnn_l2_kernel(int n0, int n1, int n2, const float *__restrict__ m,
                                                     const float *__restrict__ W2, const float *__restrict__ P,
                                                     float *__restrict__ out)
{
    __shared__ float w[NTAP * C * C];
    for (int i = threadIdx.x; i < NTAP * C * C; i += blockDim.x) w[i] = W2[i];
    __syncthreads();
    const long N = (long)n0 * n1 * n2;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= 2 * N) return;
    const int comp = (int)(I / N);
    const long v = I % N;
    const int i = (int)(v % n0), j = (int)((v / n0) % n1), k = (int)(v / ((long)n0 * n1));
    float o[C];
#pragma unroll
    for (int b = 0; b < C; ++b) o[b] = 0.f;
    for (int t = 0; t < NTAP; ++t) {
        const long q = nb(i, j, k, t, n0, n1, n2);
        if (q < 0) continue;
        const float *mq = m + ((long)comp * N + q) * C;
        for (int a = 0; a < C; ++a) {
            const float x = mq[a];
#pragma unroll
            for (int b = 0; b < C; ++b) o[b] += w[(t * C + a) * C + b] * x;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < C; ++b) s += P[v * C + b] * o[b];
    out[I] = s;
}

// ---------------------------------------------------------------- FP32 backward
// do[(comp N + v) C + b] = P[v][b] * 2 (out - target) / T2
__global__ void cnn_dout_kernel(long N, const float *__restrict__ out, const fcomplex *__restrict__ target,
                                const float *__restrict__ P, float inv_t2, float *__restrict__ dov)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= 2 * N) return;
    const int comp = (int)(I / N);
    const long v = I % N;
    const float d = 2.f * (out[I] - comp_of(target[v], comp)) * inv_t2;
#pragma unroll
    for (int b = 0; b < C; ++b) dov[I * C + b] = P[v * C + b] * d;
}

// dh[(comp N + v) C + a] = P[v][a] sum_t sum_b W2[t][a][b] do[comp][v - t][b]
__global__ void __launch_bounds__(256) cnn_dm_kernel(int n0, int n1, int n2, const float *__restrict__ dov,
                                                     const float *__restrict__ W2, const float *__restrict__ P,
                                                     float *__restrict__ dh)
{
    __shared__ float w[NTAP * C * C];
    for (int i = threadIdx.x; i < NTAP * C * C; i += blockDim.x) w[i] = W2[i];
    __syncthreads();
    const long N = (long)n0 * n1 * n2;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= 2 * N) return;
    const int comp = (int)(I / N);
    const long v = I % N;
    const int i = (int)(v % n0), j = (int)((v / n0) % n1), k = (int)(v / ((long)n0 * n1));
    float g[C];
#pragma unroll
    for (int a = 0; a < C; ++a) g[a] = 0.f;
    for (int t = 0; t < NTAP; ++t) {
        // m(v) feeds o(v - t) through tap t
        const int a0 = i - tap_dx(t), b0 = j - tap_dy(t), c0 = k - tap_dz(t);
        if (a0 < 0 || a0 >= n0 || b0 < 0 || b0 >= n1 || c0 < 0 || c0 >= n2) continue;
        const long q = a0 + (long)n0 * (b0 + (long)n1 * c0);
        const float *dq = dov + ((long)comp * N + q) * C;
        for (int b = 0; b < C; ++b) {
            const float x = dq[b];
#pragma unroll
            for (int a = 0; a < C; ++a) g[a] += w[(t * C + a) * C + b] * x;
        }
    }
#pragma unroll
    for (int a = 0; a < C; ++a) dh[I * C + a] = P[v * C + a] * g[a];
}

// Weight gradients as partial sums per block: block (t, chunk) accumulates over its voxel chunk.
// dW2[t][a][b] = sum_{comp, v} m[comp][v + t][a] do[comp][v][b]     (256 threads = (a, b))
__global__ void __launch_bounds__(256) cnn_dw2_kernel(int n0, int n1, int n2, int nchunk,
                                                      const float *__restrict__ m, const float *__restrict__ dov,
                                                      float *__restrict__ partial)
{
    const int t = blockIdx.x, chunk = blockIdx.y;
    const int a = threadIdx.x / C, b = threadIdx.x % C;
    const long N = (long)n0 * n1 * n2;
    const long per = (2 * N + nchunk - 1) / nchunk;
    const long I0 = (long)chunk * per, I1 = I0 + per < 2 * N ? I0 + per : 2 * N;
    float s = 0.f;
    for (long I = I0; I < I1; ++I) {
        const int comp = (int)(I / N);
        const long v = I % N;
        const int i = (int)(v % n0), j = (int)((v / n0) % n1), k = (int)(v / ((long)n0 * n1));
        const long q = nb(i, j, k, t, n0, n1, n2);
        if (q < 0) continue;
        s += m[((long)comp * N + q) * C + a] * dov[I * C + b];
    }
    partial[((long)chunk * NTAP + t) * C * C + a * C + b] = s;
}

// dW1[t][c][a] = sum_{comp, v} u_c,comp(v + t) dh[comp][v][a]   (block (chunk, c): C threads per tap)
__global__ void __launch_bounds__(NTAP * C) cnn_dw1_kernel(int n0, int n1, int n2, int nchunk,
                                                           const fcomplex *__restrict__ u,
                                                           const float *__restrict__ dh, float *__restrict__ partial)
{
    const int t = threadIdx.x / C, a = threadIdx.x % C, chunk = blockIdx.x, c = blockIdx.y;
    const long N = (long)n0 * n1 * n2;
    const long per = (2 * N + nchunk - 1) / nchunk;
    const long I0 = (long)chunk * per, I1 = I0 + per < 2 * N ? I0 + per : 2 * N;
    float s = 0.f;
    for (long I = I0; I < I1; ++I) {
        const int comp = (int)(I / N);
        const long v = I % N;
        const int i = (int)(v % n0), j = (int)((v / n0) % n1), k = (int)(v / ((long)n0 * n1));
        const long q = nb(i, j, k, t, n0, n1, n2);
        if (q < 0) continue;
        s += comp_of(u[(long)c * N + q], comp) * dh[I * C + a];
    }
    partial[((long)chunk * NTAP + t) * CI * C + c * C + a] = s;
}

// out[i] = sum_c partial[c * n + i] (fixed order)
__global__ void sum_partials_kernel(const float *__restrict__ partial, int nchunk, int n, float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[(long)c * n + i];
    out[i] = (float)s;
}

// ---------------------------------------------------------------- BF16 MFMA forward
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (short)(u >> 16);
}

// Layer 1 on MFMA: one wave per 16 consecutive voxels along x (one component).
// A[i][k] = u_c(v_i + tap t), k = t CI + c (zero for k >= 27 CI), B[k][a] = W1[t][c][a];
// D[i][a] -> m = P[v_i][a] D[i][a] stored as bf16 at mb[(comp N + v) C + a].
// Operand layout of v_mfma_f32_16x16x16_bf16 (wave64): lane l holds A[l%16][4(l/16)+q],
// B[4(l/16)+q][l%16], q < 4, and D[4(l/16)+q][l%16].
__global__ void __launch_bounds__(256) cnn_l1_mfma_kernel(int n0, int n1, int n2, long ntiles,
                                                          const fcomplex *__restrict__ u,
                                                          const float *__restrict__ W1, const float *__restrict__ P,
                                                          unsigned short *__restrict__ mb)
{
    static_assert(CI == 4, "a lane's 4 consecutive k must be the CI channels of one tap");
    __shared__ short wb[K1 * C];   // [k][a], k padded to K1
    for (int e = threadIdx.x; e < K1 * C; e += blockDim.x) wb[e] = e < NTAP * CI * C ? to_bf16(W1[e]) : (short)0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long tile = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    const long N = (long)n0 * n1 * n2;
    const int xt = (n0 + 15) / 16;
    const long lines = (long)n1 * n2;
    const int comp = (int)(tile / (lines * xt));
    const long rem = tile % (lines * xt);
    const long line = rem / xt;
    const int x0 = (int)(rem % xt) * 16;
    const int j = (int)(line % n1), k = (int)(line / n1);
    const int row = lane & 15, kb = (lane >> 4) * 4;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < K1; kk += 16) {
        bf16x4 a, b;
        const int t = (kk + kb) / CI;   // this lane's tap; q runs over its CI channels
        const int i = x0 + row;
        long idx = -1;
        if (t < NTAP && i < n0) idx = nb(i, j, k, t, n0, n1, n2);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[q] = to_bf16(idx >= 0 ? comp_of(u[(long)q * N + idx], comp) : 0.f);
            b[q] = wb[(kk + kb + q) * C + row];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    // lane holds D[kb + q][row]: voxel x0 + kb + q, channel row
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = x0 + kb + q;
        if (i >= n0) continue;
        const long v = i + (long)n0 * (j + (long)n1 * k);
        mb[((long)comp * N + v) * C + row] = (unsigned short)to_bf16(P[v * C + row] * acc[q]);
    }
}

// Layer 2 on MFMA: per tap t, A[i][a] = m[v_i + t][a], B[a][b] = W2[t][a][b]; the 27
// products accumulate into D[i][b]; out(v_i) = sum_b P[v_i][b] D[i][b] (reduced over
// the 16 lanes that hold row i's columns).
__global__ void __launch_bounds__(256) cnn_l2_mfma_kernel(int n0, int n1, int n2, long ntiles,
                                                          const unsigned short *__restrict__ mb,
                                                          const float *__restrict__ W2, const float *__restrict__ P,
                                                          float *__restrict__ out)
{
    __shared__ short wb[NTAP * C * C];   // [t][a][b]
    for (int e = threadIdx.x; e < NTAP * C * C; e += blockDim.x) wb[e] = to_bf16(W2[e]);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long tile = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    const long N = (long)n0 * n1 * n2;
    const int xt = (n0 + 15) / 16;
    const long lines = (long)n1 * n2;
    const int comp = (int)(tile / (lines * xt));
    const long rem = tile % (lines * xt);
    const long line = rem / xt;
    const int x0 = (int)(rem % xt) * 16;
    const int j = (int)(line % n1), k = (int)(line / n1);
    const int row = lane & 15, kb = (lane >> 4) * 4;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < NTAP; ++t) {
        bf16x4 a, b;
        const int i = x0 + row;
        long idx = -1;
        if (i < n0) idx = nb(i, j, k, t, n0, n1, n2);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[q] = idx >= 0 ? (short)mb[((long)comp * N + idx) * C + kb + q] : (short)0;
            b[q] = wb[(t * C + kb + q) * C + row];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    // lane holds D[kb + q][row] (voxel x0 + kb + q, output channel row)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = x0 + kb + q;
        float s = 0.f;
        if (i < n0) {
            const long v = i + (long)n0 * (j + (long)n1 * k);
            s = P[v * C + row] * acc[q];
        }
        // sum over the 16 lanes of this group (channels)
        for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
        if (row == 0 && i < n0) {
            const long v = i + (long)n0 * (j + (long)n1 * k);
            out[(long)comp * N + v] = s;
        }
    }
}

// n[v] = g * (out_re[v] + i out_im[v]) as complex64
__global__ void cnn_to_complex_kernel(long N, const float *__restrict__ out, fcomplex g, fcomplex *__restrict__ n)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const float re = out[v], im = out[N + v];
    n[v] = {g.re * re - g.im * im, g.re * im + g.im * re};
}

// u[c N + v] *= s.v[c] for the CI input fields
__global__ void scale_inputs_kernel(long N, CnnScales s, fcomplex *__restrict__ u)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= CI * N) return;
    const float f = s.v[I / N];
    u[I] = {u[I].re * f, u[I].im * f};
}

// first-moment features: G[v] -= x_d(v) My[v], where G holds X_d y on entry
__global__ void moment_feature_kernel(int n0, int n1, int n2, int d, const float *__restrict__ gr,
                                      const fcomplex *__restrict__ My, fcomplex *__restrict__ G)
{
    const long N = (long)n0 * n1 * n2;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int ii = d == 0 ? (int)(v % n0) : d == 1 ? (int)((v / n0) % n1) : (int)(v / ((long)n0 * n1));
    const float x = gr[ii];
    G[v] = {G[v].re - x * My[v].re, G[v].im - x * My[v].im};
}

// t = pv - Vg .* My: the part of the potential product the lumped form misses
__global__ void lumped_residual_kernel(long N, const fcomplex *__restrict__ pv, const fcomplex *__restrict__ My,
                                       const float *__restrict__ Vg, fcomplex *__restrict__ t)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const float w = Vg[v];
    t[v] = {pv[v].re - w * My[v].re, pv[v].im - w * My[v].im};
}

inline unsigned gridn(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void scale_inputs(hipStream_t st, long N, const CnnScales &s, fcomplex *u)
{
    hipLaunchKernelGGL(scale_inputs_kernel, dim3(gridn(CI * N, 256)), dim3(256), 0, st, N, s, u);
}

void moment_feature(hipStream_t st, const int n[3], int d, const float *gr, const fcomplex *My, fcomplex *G)
{
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(moment_feature_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, n[0], n[1], n[2], d, gr, My, G);
}

void lumped_residual(hipStream_t st, long N, const fcomplex *pv, const fcomplex *My, const float *Vg, fcomplex *t)
{
    hipLaunchKernelGGL(lumped_residual_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, N, pv, My, Vg, t);
}

void cnn_forward_f32(hipStream_t st, const int n[3], const fcomplex *u, const float *W1, const float *W2,
                     const float *P, float *m, float *out)
{
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(cnn_l1_kernel, dim3(gridn(2 * N, 256)), dim3(256), 0, st, n[0], n[1], n[2], u, W1, P, m);
    hipLaunchKernelGGL(cnn_l2_kernel, dim3(gridn(2 * N, 256)), dim3(256), 0, st, n[0], n[1], n[2], m, W2, P, out);
}

void cnn_backward_f32(hipStream_t st, const int n[3], const fcomplex *u, const fcomplex *target, float inv_t2,
                      const float *W2, const float *P, const float *m, const float *out, float *dov, float *dh,
                      float *partial, int nchunk, float *dW1, float *dW2)
{
    const long N = (long)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(cnn_dout_kernel, dim3(gridn(2 * N, 256)), dim3(256), 0, st, N, out, target, P, inv_t2, dov);
    hipLaunchKernelGGL(cnn_dw2_kernel, dim3(NTAP, nchunk), dim3(256), 0, st, n[0], n[1], n[2], nchunk, m, dov,
                       partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(gridn(NTAP * C * C, 256)), dim3(256), 0, st, partial, nchunk,
                       NTAP * C * C, dW2);
    hipLaunchKernelGGL(cnn_dm_kernel, dim3(gridn(2 * N, 256)), dim3(256), 0, st, n[0], n[1], n[2], dov, W2, P, dh);
    hipLaunchKernelGGL(cnn_dw1_kernel, dim3(nchunk, CI), dim3(NTAP * C), 0, st, n[0], n[1], n[2], nchunk, u, dh,
                       partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(gridn(NTAP * CI * C, 256)), dim3(256), 0, st, partial, nchunk,
                       NTAP * CI * C, dW1);
}

void cnn_infer_mfma(hipStream_t st, const int n[3], const fcomplex *u, const float *W1, const float *W2,
                    const float *P, unsigned short *mb, float *out, fcomplex g, fcomplex *nout)
{
    const long N = (long)n[0] * n[1] * n[2];
    const long ntiles = 2L * ((n[0] + 15) / 16) * n[1] * n[2];
    hipLaunchKernelGGL(cnn_l1_mfma_kernel, dim3(gridn(ntiles, 4)), dim3(256), 0, st, n[0], n[1], n[2], ntiles, u,
                       W1, P, mb);
    hipLaunchKernelGGL(cnn_l2_mfma_kernel, dim3(gridn(ntiles, 4)), dim3(256), 0, st, n[0], n[1], n[2], ntiles, mb,
                       W2, P, out);
    if (nout) hipLaunchKernelGGL(cnn_to_complex_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, N, out, g, nout);
}

void cnn_to_complex(hipStream_t st, long N, const float *out, fcomplex g, fcomplex *n)
{
    hipLaunchKernelGGL(cnn_to_complex_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, N, out, g, n);
}

}  // namespace tdz
