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

#include <algorithm>
#include <vector>

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

// Weight gradients on MFMA (v_mfma_f32_16x16x4_f32, FP32 in and out). Operand layout
// (wave64): lane l holds A[l%16][l/16], B[l/16][l%16] and D[4(l/16)+q][l%16], q < 4.
// One MFMA consumes 4 consecutive entries I of the flattened [comp][v] index (k = l/16).
// Block b covers a contiguous range of such groups, its 4 waves take every 4th group;
// the waves' accumulators are summed through LDS in wave order and the block writes
// its partial sums, reduced over the blocks afterwards in a fixed order (deterministic).
typedef float floatx4g __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void voxel_of(long I, long N, int n0, int n1, int &comp, int &i, int &j, int &k)
{
    comp = (int)(I / N);
    const long v = I % N;
    i = (int)(v % n0);
    j = (int)((v / n0) % n1);
    k = (int)(v / ((long)n0 * n1));
}

// dW2[t][a][b] = sum_I m[comp][v + t][a] do[I][b]: A[a][k] = m(v_k + t)[a], B[k][b] = do[I_k][b]
__global__ void __launch_bounds__(256) cnn_dw2_mfma_kernel(int n0, int n1, int n2, const float *__restrict__ m,
                                                           const float *__restrict__ dov,
                                                           float *__restrict__ partial)
{
    __shared__ float red[NTAP * C * C];
    const long N = (long)n0 * n1 * n2;
    const long groups = (2 * N + 3) / 4;
    const long per = (groups + gridDim.x - 1) / gridDim.x;
    const long g0 = (long)blockIdx.x * per, g1 = g0 + per < groups ? g0 + per : groups;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int col = lane & 15, kk = lane >> 4;
    floatx4g acc[NTAP];
#pragma unroll
    for (int t = 0; t < NTAP; ++t) acc[t] = {0.f, 0.f, 0.f, 0.f};
    for (long g = g0 + w; g < g1; g += 4) {
        const long I = 4 * g + kk;
        float b = 0.f;
        int comp = 0, i = 0, j = 0, k = 0;
        const bool in = I < 2 * N;
        if (in) {
            b = dov[I * C + col];
            voxel_of(I, N, n0, n1, comp, i, j, k);
        }
#pragma unroll
        for (int t = 0; t < NTAP; ++t) {
            float a = 0.f;
            if (in) {
                const long q = nb(i, j, k, t, n0, n1, n2);
                if (q >= 0) a = m[((long)comp * N + q) * C + col];
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
        }
    }
    // lane holds D[4 kk + q][col] = (a = 4 kk + q, b = col)
    for (int ww = 0; ww < 4; ++ww) {
        if (w == ww)
#pragma unroll
            for (int t = 0; t < NTAP; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float &r = red[(t * C + 4 * kk + q) * C + col];
                    r = ww == 0 ? acc[t][q] : r + acc[t][q];
                }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < NTAP * C * C; e += blockDim.x) partial[(long)blockIdx.x * NTAP * C * C + e] = red[e];
}

// dW1[t][c][a] = sum_I u_c,comp(v + t) dh[I][a]. Rows of A pair 4 taps with the CI = 4
// channels: A[(t', c)][k] = u_c(v_k + tap 4 s + t'), 7 steps s cover the 27 taps.
__global__ void __launch_bounds__(256) cnn_dw1_mfma_kernel(int n0, int n1, int n2, const fcomplex *__restrict__ u,
                                                           const float *__restrict__ dh,
                                                           float *__restrict__ partial)
{
    static_assert(CI == 4, "4 taps x CI channels fill the 16 rows of A");
    constexpr int S = (NTAP + 3) / 4;
    __shared__ float red[NTAP * CI * C];
    const long N = (long)n0 * n1 * n2;
    const long groups = (2 * N + 3) / 4;
    const long per = (groups + gridDim.x - 1) / gridDim.x;
    const long g0 = (long)blockIdx.x * per, g1 = g0 + per < groups ? g0 + per : groups;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int col = lane & 15, kk = lane >> 4;
    const int tq = col >> 2, c = col & 3;   // this lane's A row: tap offset t' and channel
    floatx4g acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = {0.f, 0.f, 0.f, 0.f};
    for (long g = g0 + w; g < g1; g += 4) {
        const long I = 4 * g + kk;
        float b = 0.f;
        int comp = 0, i = 0, j = 0, k = 0;
        const bool in = I < 2 * N;
        if (in) {
            b = dh[I * C + col];
            voxel_of(I, N, n0, n1, comp, i, j, k);
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int t = 4 * s + tq;
            float a = 0.f;
            if (in && t < NTAP) {
                const long q = nb(i, j, k, t, n0, n1, n2);
                if (q >= 0) a = comp_of(u[(long)c * N + q], comp);
            }
            acc[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[s], 0, 0, 0);
        }
    }
    // lane holds D[4 kk + q][col] = (tap 4 s + kk, channel q, hidden channel col)
    for (int ww = 0; ww < 4; ++ww) {
        if (w == ww)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int t = 4 * s + kk;
                if (t >= NTAP) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float &r = red[(t * CI + q) * C + col];
                    r = ww == 0 ? acc[s][q] : r + acc[s][q];
                }
            }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < NTAP * CI * C; e += blockDim.x) partial[(long)blockIdx.x * NTAP * CI * C + e] = red[e];
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

// round to nearest even: one v_cvt_pk_bf16_f32 on gfx950
__device__ __forceinline__ short to_bf16(float f) { return __builtin_bit_cast(short, (__bf16)f); }

// Tiles of the MFMA inference: a block of 256 threads (4 waves) owns 16 (x) x 4 (y)
// x 12 (z) voxels of one component; wave w owns z-planes 3w .. 3w+2 of the tile, each
// as four 16-voxel rows along x. The block first stages the tile plus a one-voxel halo
// (18 x 6 x 14) in LDS, so every input value is read from HBM/L2 about twice per tile
// volume instead of 27 times, and the bf16 weights (packed once per call in the lanes'
// operand order) are loaded once per block.
constexpr int TX = 16, TY = 4, TZ = 12, WZ = TZ / 4;
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
// packed weights (behind m in the work buffer): the layer-1 K steps as 8-byte operands,
// then from byte kPack2Offset the layer-2 tap pairs as 16-byte operands (2 per (dz, dy))
constexpr int NPACK1 = K1 / 16, NPACK2 = 2 * 9, NPACK = NPACK1 + NPACK2;
constexpr int kPack2Offset = 4096;
static_assert(NPACK1 * 64 * 8 <= kPack2Offset, "layer-1 slots fit before the layer-2 slots");

struct TileIdx {
    int comp, x0, y0, z0;
};

// blockIdx -> tile, XCD-aware: consecutive dispatches go round-robin over the 8 XCDs,
// so block b is remapped to tile (b % 8) * (nb / 8) + b / 8; neighbouring tiles (which
// share halo voxels) then land on the same XCD and its L2.
__device__ __forceinline__ TileIdx tile_of(int n0, int n1, int n2)
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = nb / 8;
    const int t = b < 8 * per ? (b % 8) * per + b / 8 : b;
    const int xt = (n0 + TX - 1) / TX, yt = (n1 + TY - 1) / TY, zt = (n2 + TZ - 1) / TZ;
    TileIdx r;
    r.x0 = (t % xt) * TX;
    r.y0 = ((t / xt) % yt) * TY;
    r.z0 = ((t / (xt * yt)) % zt) * TZ;
    r.comp = t / (xt * yt * zt);
    return r;
}

__device__ __forceinline__ int tap_of_halo(int lx, int ly, int lz) { return (lz * HY + ly) * HX + lx; }

// the 4 bf16 of an 8-byte word as an MFMA operand (element 0 in the low half of .x)
__device__ __forceinline__ bf16x4 unpack4(uint2 raw) { return __builtin_bit_cast(bf16x4, raw); }

__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d)
{
    return {(unsigned)(unsigned short)to_bf16(a) | ((unsigned)(unsigned short)to_bf16(b) << 16),
            (unsigned)(unsigned short)to_bf16(c) | ((unsigned)(unsigned short)to_bf16(d) << 16)};
}

// Layer 1, wp[s * 64 + lane] (s < NPACK1): the lane's 4 bf16 A-operand values W1[t][q][row],
// t = 4 s + kb / 4 (zero past the last tap).
// Layer 2, wp2[(2 p + half) * 64 + lane], p = 3 (dz + 1) + (dy + 1): the lane's 8 bf16
// A-operand values of a 16x16x32 product that takes two taps of the same (dz, dy), the
// lane group g = lane / 16 holding input channels 8 (g & 1) .. +7 of tap
//   half 0: dx = -1 (g < 2), dx = 0 (g >= 2);   half 1: dx = +1 (g < 2), zero (g >= 2),
// value W2[t][a][row].
__global__ void cnn_pack_kernel(const float *__restrict__ W1, const float *__restrict__ W2, uint2 *__restrict__ wp,
                                uint4 *__restrict__ wp2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NPACK * 64) return;
    const int slot = e / 64, lane = e % 64, row = lane & 15, g = lane >> 4;
    if (slot < NPACK1) {
        const int t = 4 * slot + g;
        float v[4];
        for (int q = 0; q < 4; ++q) v[q] = t < NTAP ? W1[(t * CI + q) * C + row] : 0.f;
        wp[slot * 64 + lane] = pack4(v[0], v[1], v[2], v[3]);
        return;
    }
    const int s2 = slot - NPACK1, p = s2 / 2, half = s2 % 2;
    const int dx = half == 0 ? (g < 2 ? -1 : 0) : 1;
    const bool zero = half == 1 && g >= 2;
    const int t = 9 * (p / 3) + 3 * (p % 3) + dx + 1;
    float v[8];
    for (int j = 0; j < 8; ++j) v[j] = zero ? 0.f : W2[(t * C + 8 * (g & 1) + j) * C + row];
    const uint2 lo = pack4(v[0], v[1], v[2], v[3]), hi = pack4(v[4], v[5], v[6], v[7]);
    wp2[s2 * 64 + lane] = {lo.x, lo.y, hi.x, hi.y};
}

// The inference evaluates the positional fields where it uses them, in FP32, from the
// Greville coordinates (the closed forms of CnnFieldGen) instead of reading a [N][C] table:
// a lane needs the fields of its channel group kb .. kb + 3 at its output voxels, about 40
// VALU operations and one reciprocal square root per voxel against 8 bytes of HBM traffic
// per voxel and lane group (222 MB per inference at 132^3). Coordinates past the grid's
// end are clamped (those voxels are not stored).
// The lane's weights of the 16 fields: the scale of field 4 k + q where k is the lane's
// channel group, zero for the other groups, so that a field value is a weighted sum with
// no lane-dependent branch (a select per field lets the compiler branch per field).
__device__ __forceinline__ void field_weights(const CnnFieldGen &f, int kb, float cw[16])
{
    const int g = kb >> 2;
#pragma unroll
    for (int e = 0; e < 16; ++e) cw[e] = g == e / 4 ? f.scale[e] : 0.f;
}
__device__ __forceinline__ float field_coord(const CnnFieldGen &f, int d, int i, int n)
{
    return f.gr[d][i < n ? i : n - 1];
}
__device__ __forceinline__ void gen_fields(const CnnFieldGen &f, const float cw[16], float x, float y, float z,
                                           float pv[4])
{
    const float r2 = x * x + y * y + z * z;
    const float rho = rsqrtf(r2 + f.a), rho2 = rho * rho, rho3 = rho2 * rho, rho5 = rho3 * rho2;
    const float V = -f.sZ * rho, s3 = f.sZ * rho3, s5 = -3.f * f.sZ * rho5;
    const float hx = x * f.h[0], hy = y * f.h[1], hz = z * f.h[2];   // x_d h_d
    const float v[16] = {1.f,
                         V,
                         s3 * hx,
                         s3 * hy,
                         s3 * hz,
                         s3 * (f.h[0] * f.h[0]) + s5 * hx * hx,
                         s3 * (f.h[1] * f.h[1]) + s5 * hy * hy,
                         s3 * (f.h[2] * f.h[2]) + s5 * hz * hz,
                         s5 * hx * hy,
                         s5 * hx * hz,
                         s5 * hy * hz,
                         V * V,
                         x * f.inv_l[0],
                         y * f.inv_l[1],
                         z * f.inv_l[2],
                         r2 * f.inv_r2};
#pragma unroll
    for (int q = 0; q < 4; ++q) pv[q] = cw[q] * v[q] + cw[4 + q] * v[4 + q] + cw[8 + q] * v[8 + q] + cw[12 + q] * v[12 + q];
}

// Layer 1 on MFMA, v_mfma_f32_16x16x16_bf16, both components per block (the inputs are
// read once). Operand layout (wave64): lane l holds A[l%16][4(l/16)+q], B[4(l/16)+q][l%16],
// q < 4, and D[4(l/16)+q][l%16]. The weights are the A operand and the data the B
// operand, so D is the transposed product: with k = t CI + c (zero for k >= 27 CI),
// A[a][k] = W1[t][c][a], B[k][i] = u_c(v_i + tap t), D[a][i] = sum_k W1 u, and lane l
// holds the 4 consecutive hidden channels a = 4(l/16) .. +3 of voxel v_{l%16}:
// m = P[v_i][a] D[a][i] is one 8-byte bf16 store at mb[(comp N + v) C + a].
// The halo tile is staged in LDS as bf16 (the CI channels of a voxel are one 8-byte
// operand, no conversion per MFMA). At least three waves per SIMD (the field evaluation
// otherwise takes the kernel to 176 registers and two).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) cnn_l1_mfma_kernel(int n0, int n1, int n2, const fcomplex *__restrict__ u,
                                                          const uint2 *__restrict__ wp, CnnFieldGen fg,
                                                          unsigned short *__restrict__ mb)
{
    static_assert(CI == 4, "a lane's 4 consecutive k must be the CI channels of one tap");
    __shared__ uint2 usv[2][HV];   // [comp][voxel]: the CI channels as bf16
    const TileIdx T = tile_of(n0, n1, n2);
    const long N = (long)n0 * n1 * n2;
    // stage one halo voxel (its CI channels) per thread and round, in two phases (all
    // loads in flight, then all LDS stores)
    constexpr int NV = (HV + 255) / 256;
    fcomplex val[NV][CI];
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int h = threadIdx.x + 256 * r;
        const int gx = T.x0 - 1 + h % HX, gy = T.y0 - 1 + (h / HX) % HY, gz = T.z0 - 1 + h / (HX * HY);
        const bool in = h < HV && gx >= 0 && gx < n0 && gy >= 0 && gy < n1 && gz >= 0 && gz < n2;
        const fcomplex *src = u + (in ? gx + (long)n0 * (gy + (long)n1 * gz) : 0);
#pragma unroll
        for (int c = 0; c < CI; ++c) val[r][c] = in ? src[(long)c * N] : fcomplex{0.f, 0.f};
    }
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int h = threadIdx.x + 256 * r;
        if (h < HV) {
            usv[0][h] = pack4(val[r][0].re, val[r][1].re, val[r][2].re, val[r][3].re);
            usv[1][h] = pack4(val[r][0].im, val[r][1].im, val[r][2].im, val[r][3].im);
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = lane & 15, kb = (lane >> 4) * 4;
    bf16x4 bw[K1 / 16];
#pragma unroll
    for (int s = 0; s < K1 / 16; ++s) bw[s] = unpack4(wp[s * 64 + lane]);
    __syncthreads();
    const uint2 *us0 = usv[0], *us1 = usv[1];
    const int gx = T.x0 + row;
    // the coordinates of this lane's outputs and its channel group's field scales
    float cw[16], fy[TY];
    field_weights(fg, kb, cw);
    const float fx = field_coord(fg, 0, gx, n0);
#pragma unroll
    for (int ry = 0; ry < TY; ++ry) fy[ry] = field_coord(fg, 1, T.y0 + ry, n1);
#pragma unroll 1   // unrolled, the plane loop runs slower (code size and register pressure)
    for (int pz = 0; pz < WZ; ++pz) {
        const int lz = w * WZ + pz, gz = T.z0 + lz;
        const float fz = field_coord(fg, 2, gz, n2);
        float pv[TY][4];
#pragma unroll
        for (int ry = 0; ry < TY; ++ry) gen_fields(fg, cw, fx, fy[ry], fz, pv[ry]);
        floatx4 acc[2][TY];
#pragma unroll
        for (int ry = 0; ry < TY; ++ry) acc[0][ry] = acc[1][ry] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < K1 / 16; ++s) {
            const int t = 4 * s + kb / 4;
            const int tt = t < NTAP ? t : 13;   // padded k: any in-tile voxel, the weight is zero there
            const int dx = tap_dx(tt), dy = tap_dy(tt), dz = tap_dz(tt);
#pragma unroll
            for (int ry = 0; ry < TY; ++ry) {
                const int h = tap_of_halo(row + 1 + dx, ry + 1 + dy, lz + 1 + dz);
                acc[0][ry] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bw[s], unpack4(us0[h]), acc[0][ry], 0, 0, 0);
                acc[1][ry] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bw[s], unpack4(us1[h]), acc[1][ry], 0, 0, 0);
            }
        }
        // lane holds D[kb + q][row]: hidden channel kb + q of voxel (x0 + row, y0 + ry)
        if (gx < n0 && gz < n2) {
#pragma unroll
            for (int ry = 0; ry < TY; ++ry) {
                const int gy = T.y0 + ry;
                if (gy >= n1) continue;
                const long v = gx + (long)n0 * (gy + (long)n1 * gz);
                const float *p = pv[ry];
                const floatx4 a0 = acc[0][ry], a1 = acc[1][ry];
                *reinterpret_cast<uint2 *>(mb + v * C + kb) =
                    pack4(p[0] * a0[0], p[1] * a0[1], p[2] * a0[2], p[3] * a0[3]);
                *reinterpret_cast<uint2 *>(mb + (N + v) * C + kb) =
                    pack4(p[0] * a1[0], p[1] * a1[1], p[2] * a1[2], p[3] * a1[3]);
            }
        }
    }
}

// Layer 2 on MFMA, v_mfma_f32_16x16x32_bf16 (lane l holds A[l%16][8(l/16)+j],
// B[8(l/16)+j][l%16], j < 8, and D[4(l/16)+q][l%16]): per (dz, dy) two products of K = 32,
// each taking two taps of the same output row (dx = -1 and 0, then dx = +1 and zero
// weights): A[b][k] = W2[t(k)][a(k)][b], B[k][i] = m[v_i + t(k)][a(k)], k < 16 the first
// tap. The 27 taps accumulate into D[b][i] (lane l holds output channels 4(l/16) .. +3 of
// voxel v_{l%16}); out(v_i) = sum_b P[v_i][b] D[b][i] (4 channels in the lane, then over
// the 4 lanes that hold voxel i's channel groups). The halo tile of m lives in LDS with a
// 48-byte voxel stride (16 bf16 channels + 8 padding): a lane's 8 channels are one aligned
// 16-byte read, and the 16 lanes of a lane group start on distinct 4-bank groups.
// Two waves per SIMD is what the 72.6 KB tile allows (two blocks per CU); the attribute
// keeps the registers within that (unconstrained, the compiler takes 260 and one wave).
constexpr int MS = 24;   // LDS stride of one voxel of m, in bf16 elements
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) cnn_l2_mfma_kernel(int n0, int n1, int n2,
                                                          const unsigned short *__restrict__ mb,
                                                          const uint4 *__restrict__ wp2, CnnFieldGen fg,
                                                          float *__restrict__ out)
{
    __shared__ uint4 ms[HV * MS / 8];
    const TileIdx T = tile_of(n0, n1, n2);
    const long N = (long)n0 * n1 * n2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = lane & 15, kb = (lane >> 4) * 4, g = lane >> 4;
    // the coordinates of this lane's outputs (their loads first: the fields are evaluated
    // while the halo loads below are in flight)
    float fy[TY], fz[WZ];
    const float fx = field_coord(fg, 0, T.x0 + row, n0);
#pragma unroll
    for (int ry = 0; ry < TY; ++ry) fy[ry] = field_coord(fg, 1, T.y0 + ry, n1);
#pragma unroll
    for (int pz = 0; pz < WZ; ++pz) fz[pz] = field_coord(fg, 2, T.z0 + w * WZ + pz, n2);
    // stage one halo voxel (its 16 channels, 2 x 16 bytes) per thread and round, in two
    // phases (all loads in flight, then all LDS stores)
    constexpr int NV = (HV + 255) / 256;
    uint4 val[NV][2];
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int h = threadIdx.x + 256 * r;
        const int gx = T.x0 - 1 + h % HX, gy = T.y0 - 1 + (h / HX) % HY, gz = T.z0 - 1 + h / (HX * HY);
        const bool in = h < HV && gx >= 0 && gx < n0 && gy >= 0 && gy < n1 && gz >= 0 && gz < n2;
        const uint4 *src = reinterpret_cast<const uint4 *>(
            mb + ((long)T.comp * N + (in ? gx + (long)n0 * (gy + (long)n1 * gz) : 0)) * C);
        val[r][0] = in ? src[0] : uint4{0u, 0u, 0u, 0u};
        val[r][1] = in ? src[1] : uint4{0u, 0u, 0u, 0u};
    }
    bf16x8 bw[NPACK2];
#pragma unroll
    for (int s = 0; s < NPACK2; ++s) bw[s] = __builtin_bit_cast(bf16x8, wp2[s * 64 + lane]);
    float pf[WZ][TY][4];   // the fields of this lane's channel group at its outputs
    {
        float cw[16];
        field_weights(fg, kb, cw);
#pragma unroll
        for (int pz = 0; pz < WZ; ++pz)
#pragma unroll
            for (int ry = 0; ry < TY; ++ry) gen_fields(fg, cw, fx, fy[ry], fz[pz], pf[pz][ry]);
    }
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int h = threadIdx.x + 256 * r;
        if (h < HV) {
            ms[h * (MS / 8)] = val[r][0];
            ms[h * (MS / 8) + 1] = val[r][1];
        }
    }
    __syncthreads();
    // Halo-row-major order: the two B fragments of a halo row (16 voxels shifted by dx =
    // -1 | 0 and by +1, 8 channels per lane) are read from LDS once and feed the up to 9
    // output rows (ry, pz) they are taps of.
    const int dxa = g < 2 ? -1 : 0, cq = (g & 1);   // first fragment's shift; 16-byte quarter of the voxel
    floatx4 acc[WZ][TY];
#pragma unroll
    for (int pz = 0; pz < WZ; ++pz)
#pragma unroll
        for (int ry = 0; ry < TY; ++ry) acc[pz][ry] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int hz = 0; hz < WZ + 2; ++hz)
#pragma unroll
        for (int hy = 0; hy < TY + 2; ++hy) {
            const int ha = tap_of_halo(row + 1 + dxa, hy, w * WZ + hz), hb = tap_of_halo(row + 2, hy, w * WZ + hz);
            const bf16x8 fa = __builtin_bit_cast(bf16x8, ms[ha * (MS / 8) + cq]);
            const bf16x8 fb = __builtin_bit_cast(bf16x8, ms[hb * (MS / 8) + cq]);
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz) {
                const int pz = hz - 1 - dz;   // output plane w WZ + pz reads halo plane w WZ + hz
                if (pz < 0 || pz >= WZ) continue;
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy) {
                    const int ry = hy - 1 - dy;
                    if (ry < 0 || ry >= TY) continue;
                    const int p = 3 * (dz + 1) + (dy + 1);
                    acc[pz][ry] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[2 * p], fa, acc[pz][ry], 0, 0, 0);
                    acc[pz][ry] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[2 * p + 1], fb, acc[pz][ry], 0, 0, 0);
                }
            }
        }
    static_assert(TY == 4, "the 4 channel groups of a voxel end with one row each");
    const int gx = T.x0 + row;
#pragma unroll
    for (int pz = 0; pz < WZ; ++pz) {
        const int lz = w * WZ + pz, gz = T.z0 + lz;
        // lane holds D[kb + q][row]: output channel kb + q of voxel (x0 + row, y0 + ry)
        float v4[TY];
#pragma unroll
        for (int ry = 0; ry < TY; ++ry) {
            const float *pv = pf[pz][ry];
            v4[ry] = pv[0] * acc[pz][ry][0] + pv[1] * acc[pz][ry][1] + pv[2] * acc[pz][ry][2] +
                     pv[3] * acc[pz][ry][3];
        }
        // sum over the 4 channel groups (lanes row, row + 16, row + 32, row + 48) by
        // recursive halving over ry: group g ends with the sum for ry = g (3 exchanges).
        // The lane-dependent choice is a bit mask (a select becomes a runtime index).
#pragma unroll
        for (int m = 2; m > 0; m >>= 1) {
            const unsigned um = (g & m) ? 0xFFFFFFFFu : 0u;
#pragma unroll
            for (int i = 0; i < m; ++i) {
                const unsigned lo = __float_as_uint(v4[i]), hi = __float_as_uint(v4[i + m]);
                const float keep = __uint_as_float((hi & um) | (lo & ~um));
                const float send = __uint_as_float((lo & um) | (hi & ~um));
                v4[i] = keep + __shfl_xor(send, 16 * m);
            }
        }
        const int gy = T.y0 + g;
        if (gx < n0 && gy < n1 && gz < n2) out[(long)T.comp * N + gx + (long)n0 * (gy + (long)n1 * gz)] = v4[0];
    }
}

// partial[b] = sum over this block's grid-stride share of (out - target)^2 (FP64)
__global__ void __launch_bounds__(256) cnn_loss_kernel(long N, const float *__restrict__ out,
                                                       const fcomplex *__restrict__ target,
                                                       double *__restrict__ partial)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (long I = (long)blockIdx.x * blockDim.x + threadIdx.x; I < 2 * N; I += (long)gridDim.x * blockDim.x) {
        const int comp = (int)(I / N);
        const double d = (double)out[I] - (double)comp_of(target[I % N], comp);
        s += d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// n[v] = g * (out_re[v] + i out_im[v]) as complex64
__global__ void cnn_to_complex_kernel(long N, const float *__restrict__ out, fcomplex g, fcomplex *__restrict__ n)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const float re = out[v], im = out[N + v];
    n[v] = {g.re * re - g.im * im, g.re * im + g.im * re};
}

// ---------------------------------------------------------------- fused products
// The corrected solve with the network needs, for y = S^-1 r (complex64),
//   My, corr = (Cx (x) My (x) Mz + Mx (x) Cy (x) Mz + Mx (x) My (x) Cz) y,
//   G_d = X_d y - x_d My (X_x = Xx (x) My (x) Mz etc., x_d the Greville coordinate),
// and the scaled network input u = [s0 y, s1 G_x, s2 G_y, s3 G_z]. Three passes share
// the partial products (1 -> 3 -> 4 -> 6 vectors) instead of four Kronecker applies:
//   x: P = Mx y, Q = Xx y, R = Cx y
//   y: PP = My P, PX = Xy P, QQ = My Q, S = My R + Cy P
//   z: My = Mz PP, X_x y = Mz QQ, X_y y = Mz PX, X_z y = Xz PP, corr = Mz S + Cz PP
// Bands are stored row-major [i][2p+1] (real M, X; complex C).
__global__ void feat_x_kernel(int n0, long lines, int p, const float *__restrict__ M, const float *__restrict__ X,
                              const fcomplex *__restrict__ Cb, const fcomplex *__restrict__ y,
                              fcomplex *__restrict__ P, fcomplex *__restrict__ Q, fcomplex *__restrict__ R)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= lines * n0) return;
    const int i = (int)(I % n0);
    const long base = I - i;
    const int bw = 2 * p + 1;
    const int lo = i - p < 0 ? 0 : i - p, hi = i + p > n0 - 1 ? n0 - 1 : i + p;
    float pr = 0, pi = 0, qr = 0, qi = 0, rr = 0, ri = 0;
    for (int J = lo; J <= hi; ++J) {
        const fcomplex v = y[base + J];
        const int e = i * bw + (J - i + p);
        const float m = M[e], x = X[e];
        const fcomplex cb = Cb[e];
        pr += m * v.re; pi += m * v.im;
        qr += x * v.re; qi += x * v.im;
        rr += cb.re * v.re - cb.im * v.im;
        ri += cb.re * v.im + cb.im * v.re;
    }
    P[I] = {pr, pi};
    Q[I] = {qr, qi};
    R[I] = {rr, ri};
}

__global__ void feat_y_kernel(int n0, int n1, long N, int p, const float *__restrict__ M, const float *__restrict__ X,
                              const fcomplex *__restrict__ Cb, const fcomplex *__restrict__ P,
                              const fcomplex *__restrict__ Q, const fcomplex *__restrict__ R,
                              fcomplex *__restrict__ PP, fcomplex *__restrict__ PX, fcomplex *__restrict__ QQ,
                              fcomplex *__restrict__ S)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int i = (int)(I % n0);
    const int j = (int)((I / n0) % n1);
    const long k = I / ((long)n0 * n1);
    const int bw = 2 * p + 1;
    const int lo = j - p < 0 ? 0 : j - p, hi = j + p > n1 - 1 ? n1 - 1 : j + p;
    float a0 = 0, a1 = 0, b0 = 0, b1 = 0, c0 = 0, c1 = 0, d0 = 0, d1 = 0;
    for (int J = lo; J <= hi; ++J) {
        const long src = i + (long)n0 * (J + (long)n1 * k);
        const fcomplex pv = P[src], qv = Q[src], rv = R[src];
        const int e = j * bw + (J - j + p);
        const float m = M[e], x = X[e];
        const fcomplex cb = Cb[e];
        a0 += m * pv.re; a1 += m * pv.im;
        b0 += x * pv.re; b1 += x * pv.im;
        c0 += m * qv.re; c1 += m * qv.im;
        d0 += m * rv.re + cb.re * pv.re - cb.im * pv.im;
        d1 += m * rv.im + cb.re * pv.im + cb.im * pv.re;
    }
    PP[I] = {a0, a1};
    PX[I] = {b0, b1};
    QQ[I] = {c0, c1};
    S[I] = {d0, d1};
}

// z pass; also writes u (the scaled network input, [CI][N]) when u is not null
__global__ void feat_z_kernel(int n0, int n1, int n2, int p, const float *__restrict__ M, const float *__restrict__ X,
                              const fcomplex *__restrict__ Cb, const float *__restrict__ gx,
                              const float *__restrict__ gy, const float *__restrict__ gz, CnnScales sc,
                              const fcomplex *__restrict__ y, const fcomplex *__restrict__ PP,
                              const fcomplex *__restrict__ PX, const fcomplex *__restrict__ QQ,
                              const fcomplex *__restrict__ S, fcomplex *__restrict__ My, fcomplex *__restrict__ corr,
                              fcomplex *__restrict__ u)
{
    const long plane = (long)n0 * n1, N = plane * n2;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const long mm = I % plane;
    const int k = (int)(I / plane);
    const int bw = 2 * p + 1;
    const int lo = k - p < 0 ? 0 : k - p, hi = k + p > n2 - 1 ? n2 - 1 : k + p;
    float m0 = 0, m1 = 0, xx0 = 0, xx1 = 0, xy0 = 0, xy1 = 0, xz0 = 0, xz1 = 0, c0 = 0, c1 = 0;
    for (int K = lo; K <= hi; ++K) {
        const long src = mm + plane * K;
        const fcomplex pp = PP[src], px = PX[src], qq = QQ[src], sv = S[src];
        const int e = k * bw + (K - k + p);
        const float m = M[e], x = X[e];
        const fcomplex cb = Cb[e];
        m0 += m * pp.re; m1 += m * pp.im;
        xx0 += m * qq.re; xx1 += m * qq.im;
        xy0 += m * px.re; xy1 += m * px.im;
        xz0 += x * pp.re; xz1 += x * pp.im;
        c0 += m * sv.re + cb.re * pp.re - cb.im * pp.im;
        c1 += m * sv.im + cb.re * pp.im + cb.im * pp.re;
    }
    My[I] = {m0, m1};
    corr[I] = {c0, c1};
    if (!u) return;
    const float x0 = gx[I % n0], x1 = gy[(I / n0) % n1], x2 = gz[k];
    const fcomplex yv = y[I];
    u[I] = {sc.v[0] * yv.re, sc.v[0] * yv.im};
    u[N + I] = {sc.v[1] * (xx0 - x0 * m0), sc.v[1] * (xx1 - x0 * m1)};
    u[2 * N + I] = {sc.v[2] * (xy0 - x1 * m0), sc.v[2] * (xy1 - x1 * m1)};
    u[3 * N + I] = {sc.v[3] * (xz0 - x2 * m0), sc.v[3] * (xz1 - x2 * m1)};
}

// u[c N + v] *= s.v[c] for the CI input fields
__global__ void scale_inputs_kernel(long N, CnnScales s, fcomplex *__restrict__ u)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= CI * N) return;
    const float f = s.v[I / N];
    u[I] = {u[I].re * f, u[I].im * f};
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

// A correction sweep's right-hand side in two passes (3D, equal orders P):
//   corr_xy_kernel: per z plane and 16 x 16 tile of (x, y) columns, U1 = (My (x) Mx) y and
//     U2 = (My (x) Cx + Cy (x) Mx) y (the plane with its halo staged in LDS, the x band
//     passes on the 16 + 2P rows the y passes need, then the y passes); a block takes PZ
//     consecutive planes (the next plane's loads in flight during this one's passes);
//   corr_z_kernel: per (x, y) column and chunk of z, the z band passes over a ring of the
//     last 2P + 1 input planes, q = r - g Vg (Mz U1) - (Mz U2 + Cz U1) (r null: zero).
// Every plane's x and y passes run once (the one-pass z-marching form repeats them on 2P
// halo planes per z chunk).
template <int P, int PZ>
__global__ void __launch_bounds__(256) corr_xy_kernel(int n0, int n1, int n2, FusedBands fb,
                                                      const fcomplex *__restrict__ y, fcomplex *__restrict__ U1,
                                                      fcomplex *__restrict__ U2)
{
    constexpr int R = 2 * P + 1, T = 16, H = T + 2 * P, NX = (H * H + 255) / 256;
    __shared__ fcomplex Ys[H * H];
    __shared__ fcomplex Ts[H * T], Rs[H * T];
    const int tid = threadIdx.x, ti = tid % T, tj = tid / T;
    const int tilesx = (n0 + T - 1) / T, tilesy = (n1 + T - 1) / T;
    int b = blockIdx.x;
    const int bx = b % tilesx;
    b /= tilesx;
    const int by = b % tilesy, bz = b / tilesy;
    const int i0 = bx * T, j0 = by * T, k0 = bz * PZ, k1 = min(k0 + PZ, n2);
    const int i = i0 + ti, j = j0 + tj;
    const bool own = i < n0 && j < n1;
    const int ic = min(i, n0 - 1), jc = min(j, n1 - 1);
    float mx[R], my[R];
    fcomplex cx[R], cy[R];
#pragma unroll
    for (int l = 0; l < R; ++l) {
        mx[l] = fb.M[0][ic * R + l];
        cx[l] = fb.C[0][ic * R + l];
        my[l] = fb.M[1][jc * R + l];
        cy[l] = fb.C[1][jc * R + l];
    }
    const long plane = (long)n0 * n1;
    fcomplex yv[NX];
    auto load_plane = [&](int K) {
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int e = tid + 256 * q;
            const int ii = i0 - P + e % H, jj = j0 - P + e / H;
            yv[q] = {0.f, 0.f};
            if (e < H * H && ii >= 0 && ii < n0 && jj >= 0 && jj < n1) yv[q] = y[K * plane + (long)jj * n0 + ii];
        }
    };
    load_plane(k0);
    for (int K = k0; K < k1; ++K) {
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int e = tid + 256 * q;
            if (e < H * H) Ys[e] = yv[q];
        }
        __syncthreads();
        if (K + 1 < k1) load_plane(K + 1);
        for (int e = tid; e < H * T; e += 256) {   // e % T == ti
            const int jj = e / T;
            float tr = 0, tim = 0, rr = 0, ri = 0;
#pragma unroll
            for (int l = 0; l < R; ++l) {
                const fcomplex v = Ys[jj * H + ti + l];
                tr += mx[l] * v.re;
                tim += mx[l] * v.im;
                rr += cx[l].re * v.re - cx[l].im * v.im;
                ri += cx[l].re * v.im + cx[l].im * v.re;
            }
            Ts[e] = {tr, tim};
            Rs[e] = {rr, ri};
        }
        __syncthreads();
        float p0 = 0, p1 = 0, s0 = 0, s1 = 0;
#pragma unroll
        for (int l = 0; l < R; ++l) {
            const int o = (tj + l) * T + ti;
            const fcomplex tv = Ts[o], rv = Rs[o];
            p0 += my[l] * tv.re;
            p1 += my[l] * tv.im;
            s0 += my[l] * rv.re + cy[l].re * tv.re - cy[l].im * tv.im;
            s1 += my[l] * rv.im + cy[l].re * tv.im + cy[l].im * tv.re;
        }
        if (own) {
            const long I = K * plane + (long)j * n0 + i;
            U1[I] = {p0, p1};
            U2[I] = {s0, s1};
        }
    }
}

template <int P>
__global__ void __launch_bounds__(256) corr_z_kernel(int n0, int n1, int n2, int kchunk, FusedBands fb,
                                                     const fcomplex *__restrict__ U1, const fcomplex *__restrict__ U2,
                                                     const fcomplex *__restrict__ rhs, const float *__restrict__ Vg,
                                                     fcomplex gc, fcomplex *__restrict__ q)
{
    constexpr int R = 2 * P + 1;
    const long plane = (long)n0 * n1;
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= plane) return;   // no barriers below
    const int k0 = blockIdx.y * kchunk, k1 = min(k0 + kchunk, n2);
    if (k0 >= k1) return;
    // ring of the input planes K - 2P .. K (slot K mod R); planes outside the axis are zero
    fcomplex u1[R], u2[R];
#pragma unroll
    for (int l = 0; l < R; ++l) u1[l] = u2[l] = {0.f, 0.f};
    const int Kbeg = max(k0 - P, 0), Kend = min(k1 + P, n2);
    for (int Kb = (Kbeg / R) * R; Kb < k1 + P; Kb += R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int K = Kb + r;
            if (K < Kbeg) continue;
            if (K < Kend) {
                u1[r] = U1[K * plane + col];
                u2[r] = U2[K * plane + col];
            } else {
                u1[r] = u2[r] = {0.f, 0.f};
            }
            const int ko = K - P;   // its inputs ko - P .. ko + P are in the ring
            if (ko < k0 || ko >= k1) continue;
            float m0 = 0, m1 = 0, c0 = 0, c1 = 0;
#pragma unroll
            for (int d = -P; d <= P; ++d) {
                const int kk = ko + d;   // band entry (ko, kk) = ko * R + P + d
                if (kk < 0 || kk >= n2) continue;
                const int e = ko * R + P + d;
                const float mz = fb.M[2][e];
                const fcomplex cz = fb.C[2][e];
                const int sl = (r + R + d - P) % R;   // slot of plane kk = K + d - P
                const fcomplex a = u1[sl], c = u2[sl];
                m0 += mz * a.re;
                m1 += mz * a.im;
                c0 += mz * c.re + cz.re * a.re - cz.im * a.im;
                c1 += mz * c.im + cz.re * a.im + cz.im * a.re;
            }
            const long I = ko * plane + col;
            const fcomplex rv = rhs ? rhs[I] : fcomplex{0.f, 0.f};
            const float v = Vg[I];
            q[I] = {rv.re - v * (gc.re * m0 - gc.im * m1) - c0, rv.im - v * (gc.re * m1 + gc.im * m0) - c1};
        }
    }
}

// The network sweep's products in two passes (3D, equal orders P): My, corr and the
// network's input from four x/y products:
//   feat_xy_kernel: per z plane and 16 x 16 tile of (x, y) columns, the four x/y products
//     PP = (My (x) Mx) y, QQ = (My (x) Xx) y, PX = (Xy (x) Mx) y and
//     S = (My (x) Cx + Cy (x) Mx) y (the plane with its halo in LDS; PZ planes per block);
//   feat_zr_kernel: per (x, y) column and chunk of z, the z band passes over a ring of the
//     last 2P + 1 planes of the four products: My = Mz PP, corr = Mz S + Cz PP and the
//     scaled network input u = [y, Mz QQ - x My, Mz PX - y My, Xz PP - z My].
template <int P, int PZ>
__global__ void __launch_bounds__(256) feat_xy_kernel(int n0, int n1, int n2, FusedBands fb,
                                                      const fcomplex *__restrict__ y, fcomplex *__restrict__ PPo,
                                                      fcomplex *__restrict__ QQo, fcomplex *__restrict__ PXo,
                                                      fcomplex *__restrict__ So)
{
    constexpr int R = 2 * P + 1, T = 16, H = T + 2 * P, NX = (H * H + 255) / 256;
    __shared__ fcomplex Ys[H * H];
    __shared__ fcomplex Ts[H * T], Qs[H * T], Rs[H * T];   // x passes: M, X, C
    const int tid = threadIdx.x, ti = tid % T, tj = tid / T;
    const int tilesx = (n0 + T - 1) / T, tilesy = (n1 + T - 1) / T;
    int b = blockIdx.x;
    const int bx = b % tilesx;
    b /= tilesx;
    const int by = b % tilesy, bz = b / tilesy;
    const int i0 = bx * T, j0 = by * T, k0 = bz * PZ, k1 = min(k0 + PZ, n2);
    const int i = i0 + ti, j = j0 + tj;
    const bool own = i < n0 && j < n1;
    const int ic = min(i, n0 - 1), jc = min(j, n1 - 1);
    float mx[R], xx[R], my[R], xy[R];
    fcomplex cx[R], cy[R];
#pragma unroll
    for (int l = 0; l < R; ++l) {
        mx[l] = fb.M[0][ic * R + l];
        xx[l] = fb.X[0][ic * R + l];
        cx[l] = fb.C[0][ic * R + l];
        my[l] = fb.M[1][jc * R + l];
        xy[l] = fb.X[1][jc * R + l];
        cy[l] = fb.C[1][jc * R + l];
    }
    const long plane = (long)n0 * n1;
    fcomplex yv[NX];
    auto load_plane = [&](int K) {
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int e = tid + 256 * q;
            const int ii = i0 - P + e % H, jj = j0 - P + e / H;
            yv[q] = {0.f, 0.f};
            if (e < H * H && ii >= 0 && ii < n0 && jj >= 0 && jj < n1) yv[q] = y[K * plane + (long)jj * n0 + ii];
        }
    };
    load_plane(k0);
    for (int K = k0; K < k1; ++K) {
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int e = tid + 256 * q;
            if (e < H * H) Ys[e] = yv[q];
        }
        __syncthreads();
        if (K + 1 < k1) load_plane(K + 1);
        for (int e = tid; e < H * T; e += 256) {   // e % T == ti
            const int jj = e / T;
            float tr = 0, tim = 0, qr = 0, qi = 0, rr = 0, ri = 0;
#pragma unroll
            for (int l = 0; l < R; ++l) {
                const fcomplex v = Ys[jj * H + ti + l];
                tr += mx[l] * v.re;
                tim += mx[l] * v.im;
                qr += xx[l] * v.re;
                qi += xx[l] * v.im;
                rr += cx[l].re * v.re - cx[l].im * v.im;
                ri += cx[l].re * v.im + cx[l].im * v.re;
            }
            Ts[e] = {tr, tim};
            Qs[e] = {qr, qi};
            Rs[e] = {rr, ri};
        }
        __syncthreads();
        float pp0 = 0, pp1 = 0, px0 = 0, px1 = 0, qq0 = 0, qq1 = 0, s0 = 0, s1 = 0;
#pragma unroll
        for (int l = 0; l < R; ++l) {
            const int o = (tj + l) * T + ti;
            const fcomplex tv = Ts[o], qv = Qs[o], rv = Rs[o];
            pp0 += my[l] * tv.re;
            pp1 += my[l] * tv.im;
            px0 += xy[l] * tv.re;
            px1 += xy[l] * tv.im;
            qq0 += my[l] * qv.re;
            qq1 += my[l] * qv.im;
            s0 += my[l] * rv.re + cy[l].re * tv.re - cy[l].im * tv.im;
            s1 += my[l] * rv.im + cy[l].re * tv.im + cy[l].im * tv.re;
        }
        if (own) {
            const long I = K * plane + (long)j * n0 + i;
            PPo[I] = {pp0, pp1};
            QQo[I] = {qq0, qq1};
            PXo[I] = {px0, px1};
            So[I] = {s0, s1};
        }
    }
}

template <int P>
__global__ void __launch_bounds__(256) feat_zr_kernel(int n0, int n1, int n2, int kchunk, FusedBands fb, CnnScales sc,
                                                      const fcomplex *__restrict__ y, const fcomplex *__restrict__ PP,
                                                      const fcomplex *__restrict__ QQ, const fcomplex *__restrict__ PX,
                                                      const fcomplex *__restrict__ S, fcomplex *__restrict__ My,
                                                      fcomplex *__restrict__ corr, fcomplex *__restrict__ u)
{
    constexpr int R = 2 * P + 1;
    const long plane = (long)n0 * n1, N = plane * n2;
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= plane) return;   // no barriers below
    const int k0 = blockIdx.y * kchunk, k1 = min(k0 + kchunk, n2);
    if (k0 >= k1) return;
    const float x0 = fb.g[0][col % n0], x1 = fb.g[1][col / n0];
    // ring of the input planes K - 2P .. K of the four products (slot K mod R)
    fcomplex pr[R], qr[R], xr[R], sr[R];
#pragma unroll
    for (int l = 0; l < R; ++l) pr[l] = qr[l] = xr[l] = sr[l] = {0.f, 0.f};
    const int Kbeg = max(k0 - P, 0), Kend = min(k1 + P, n2);
    for (int Kb = (Kbeg / R) * R; Kb < k1 + P; Kb += R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int K = Kb + r;
            if (K < Kbeg) continue;
            if (K < Kend) {
                const long src = K * plane + col;
                pr[r] = PP[src];
                qr[r] = QQ[src];
                xr[r] = PX[src];
                sr[r] = S[src];
            } else {
                pr[r] = qr[r] = xr[r] = sr[r] = {0.f, 0.f};
            }
            const int ko = K - P;   // its inputs ko - P .. ko + P are in the ring
            if (ko < k0 || ko >= k1) continue;
            float m0 = 0, m1 = 0, a0 = 0, a1 = 0, b0 = 0, b1 = 0, z0 = 0, z1 = 0, c0 = 0, c1 = 0;
#pragma unroll
            for (int d = -P; d <= P; ++d) {
                const int kk = ko + d;   // band entry (ko, kk) = ko * R + P + d
                if (kk < 0 || kk >= n2) continue;
                const int e = ko * R + P + d;
                const float mz = fb.M[2][e], xz = fb.X[2][e];
                const fcomplex cz = fb.C[2][e];
                const int sl = (r + R + d - P) % R;
                const fcomplex pv = pr[sl], qv = qr[sl], xv = xr[sl], sv = sr[sl];
                m0 += mz * pv.re;
                m1 += mz * pv.im;
                a0 += mz * qv.re;
                a1 += mz * qv.im;
                b0 += mz * xv.re;
                b1 += mz * xv.im;
                z0 += xz * pv.re;
                z1 += xz * pv.im;
                c0 += mz * sv.re + cz.re * pv.re - cz.im * pv.im;
                c1 += mz * sv.im + cz.re * pv.im + cz.im * pv.re;
            }
            const long I = ko * plane + col;
            My[I] = {m0, m1};
            corr[I] = {c0, c1};
            if (u) {
                const float x2 = fb.g[2][ko];
                const fcomplex yv = y[I];
                u[I] = {sc.v[0] * yv.re, sc.v[0] * yv.im};
                u[N + I] = {sc.v[1] * (a0 - x0 * m0), sc.v[1] * (a1 - x0 * m1)};
                u[2 * N + I] = {sc.v[2] * (b0 - x1 * m0), sc.v[2] * (b1 - x1 * m1)};
                u[3 * N + I] = {sc.v[3] * (z0 - x2 * m0), sc.v[3] * (z1 - x2 * m1)};
            }
        }
    }
}

inline unsigned gridn(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void fused_products(hipStream_t st, const FusedBands &fb, const CnnScales &sc, const fcomplex *y, fcomplex *tmp[7],
                    fcomplex *My, fcomplex *corr, fcomplex *u)
{
    const int n0 = fb.n[0], n1 = fb.n[1], n2 = fb.n[2];
    const long N = (long)n0 * n1 * n2;
    const int p = fb.p[0];
    if (p == fb.p[1] && p == fb.p[2] && p >= 2 && p <= 5 && n2 > 2 * p) {
        constexpr int PZ = 4;
        const unsigned nxy = (unsigned)(((n0 + 15) / 16) * ((n1 + 15) / 16) * ((n2 + PZ - 1) / PZ));
        const long plane = (long)n0 * n1;
        const unsigned cb = (unsigned)((plane + 255) / 256);
        const int chunks = std::max(1, std::min((int)(1024 / std::max(1u, cb)), n2 / (2 * p + 2)));
        const int kchunk = (n2 + chunks - 1) / chunks;
        const dim3 gz(cb, (unsigned)((n2 + kchunk - 1) / kchunk));
#define TDZ_F2(PP)                                                                                                    \
    hipLaunchKernelGGL((feat_xy_kernel<PP, PZ>), dim3(nxy), dim3(256), 0, st, n0, n1, n2, fb, y, tmp[0], tmp[1],        \
                       tmp[2], tmp[3]);                                                                             \
    hipLaunchKernelGGL((feat_zr_kernel<PP>), gz, dim3(256), 0, st, n0, n1, n2, kchunk, fb, sc, y, tmp[0], tmp[1],       \
                       tmp[2], tmp[3], My, corr, u)
        switch (p) {
        case 2: TDZ_F2(2); break;
        case 3: TDZ_F2(3); break;
        case 4: TDZ_F2(4); break;
        default: TDZ_F2(5); break;
        }
#undef TDZ_F2
        return;
    }
    hipLaunchKernelGGL(feat_x_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, n0, N / n0, fb.p[0], fb.M[0], fb.X[0],
                       fb.C[0], y, tmp[0], tmp[1], tmp[2]);
    hipLaunchKernelGGL(feat_y_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, n0, n1, N, fb.p[1], fb.M[1], fb.X[1],
                       fb.C[1], tmp[0], tmp[1], tmp[2], tmp[3], tmp[4], tmp[5], tmp[6]);
    hipLaunchKernelGGL(feat_z_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, n0, n1, n2, fb.p[2], fb.M[2], fb.X[2],
                       fb.C[2], fb.g[0], fb.g[1], fb.g[2], sc, y, tmp[3], tmp[4], tmp[5], tmp[6], My, corr, u);
}

bool fused_correction(hipStream_t st, const FusedBands &fb, const fcomplex *y, const fcomplex *r, const float *Vg,
                      fcomplex g, fcomplex *q, fcomplex *s1, fcomplex *s2)
{
    if (s1 && s2 && fb.p[0] == fb.p[1] && fb.p[0] == fb.p[2] && fb.p[0] >= 2 && fb.p[0] <= 5 &&
        fb.n[2] > 2 * fb.p[0]) {   // the two-pass form
        const int n0 = fb.n[0], n1 = fb.n[1], n2 = fb.n[2];
        constexpr int PZ = 4;
        const unsigned nxy = (unsigned)(((n0 + 15) / 16) * ((n1 + 15) / 16) * ((n2 + PZ - 1) / PZ));
        // z chunks of at least 2P + 2 planes, about a thousand blocks in all
        const long plane = (long)n0 * n1;
        const unsigned cb = (unsigned)((plane + 255) / 256);
        const int p = fb.p[0];
        const int chunks = std::max(1, std::min((int)(1024 / std::max(1u, cb)), n2 / (2 * p + 2)));
        const int kchunk = (n2 + chunks - 1) / chunks;
        const dim3 gz(cb, (unsigned)((n2 + kchunk - 1) / kchunk));
#define TDZ_C2(PP)                                                                                                 \
    hipLaunchKernelGGL((corr_xy_kernel<PP, PZ>), dim3(nxy), dim3(256), 0, st, n0, n1, n2, fb, y, s1, s2);        \
    hipLaunchKernelGGL((corr_z_kernel<PP>), gz, dim3(256), 0, st, n0, n1, n2, kchunk, fb, s1, s2, r, Vg, g, q)
        switch (p) {
        case 2: TDZ_C2(2); break;
        case 3: TDZ_C2(3); break;
        case 4: TDZ_C2(4); break;
        default: TDZ_C2(5); break;
        }
#undef TDZ_C2
        return true;
    }
    return false;
}

namespace {
// r[i] = two values uniform in [-1, 1) from a 64-bit mix (splitmix64) of (seed, i)
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void random_field_kernel(long N, unsigned long long seed, fcomplex *__restrict__ r)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        const unsigned long long h = mix64(seed * 0x100000001B3ull + (unsigned long long)i);
        const float a = (float)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;                 // 24 bits
        const float b = (float)((h >> 16) & 0xFFFFFFull) * (1.0f / 8388608.0f) - 1.0f;
        r[i] = {a, b};
    }
}
}  // namespace

void random_field(hipStream_t st, long N, unsigned long long seed, fcomplex *r)
{
    hipLaunchKernelGGL(random_field_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, N, seed, r);
}

void scale_inputs(hipStream_t st, long N, const CnnScales &s, fcomplex *u)
{
    hipLaunchKernelGGL(scale_inputs_kernel, dim3(gridn(CI * N, 256)), dim3(256), 0, st, N, s, u);
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
    hipLaunchKernelGGL(cnn_dw2_mfma_kernel, dim3(nchunk), dim3(256), 0, st, n[0], n[1], n[2], m, dov, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(gridn(NTAP * C * C, 256)), dim3(256), 0, st, partial, nchunk,
                       NTAP * C * C, dW2);
    hipLaunchKernelGGL(cnn_dm_kernel, dim3(gridn(2 * N, 256)), dim3(256), 0, st, n[0], n[1], n[2], dov, W2, P, dh);
    hipLaunchKernelGGL(cnn_dw1_mfma_kernel, dim3(nchunk), dim3(256), 0, st, n[0], n[1], n[2], u, dh, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(gridn(NTAP * CI * C, 256)), dim3(256), 0, st, partial, nchunk,
                       NTAP * CI * C, dW1);
}

void cnn_infer_mfma(hipStream_t st, const int n[3], const fcomplex *u, const float *W1, const float *W2,
                    const CnnFieldGen &fg, unsigned short *mb, float *out, fcomplex g, fcomplex *nout, bool pack)
{
    const long N = (long)n[0] * n[1] * n[2];
    const long tiles = (long)((n[0] + TX - 1) / TX) * ((n[1] + TY - 1) / TY) * ((n[2] + TZ - 1) / TZ);
    // the packed weights go behind the 2 N C bf16 of m (the work buffer holds both)
    uint2 *wp = reinterpret_cast<uint2 *>(mb + 2 * N * C);
    uint4 *wp2 = reinterpret_cast<uint4 *>(reinterpret_cast<char *>(wp) + kPack2Offset);
    if (pack) hipLaunchKernelGGL(cnn_pack_kernel, dim3(gridn(NPACK * 64, 256)), dim3(256), 0, st, W1, W2, wp, wp2);
    // layer 1: one block per tile (both components); layer 2: one per tile and component
    hipLaunchKernelGGL(cnn_l1_mfma_kernel, dim3((unsigned)tiles), dim3(256), 0, st, n[0], n[1], n[2], u, wp, fg, mb);
    hipLaunchKernelGGL(cnn_l2_mfma_kernel, dim3((unsigned)(2 * tiles)), dim3(256), 0, st, n[0], n[1], n[2], mb, wp2,
                       fg, out);
    if (nout) hipLaunchKernelGGL(cnn_to_complex_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, N, out, g, nout);
}

void cnn_loss(hipStream_t st, long N, const float *out, const fcomplex *target, double *partial)
{
    hipLaunchKernelGGL(cnn_loss_kernel, dim3(kLossBlocks), dim3(256), 0, st, N, out, target, partial);
}

void cnn_to_complex(hipStream_t st, long N, const float *out, fcomplex g, fcomplex *n)
{
    hipLaunchKernelGGL(cnn_to_complex_kernel, dim3(gridn(N, 256)), dim3(256), 0, st, N, out, g, n);
}

}  // namespace tdz
