// Per-column MLP surrogate of the LW+SW heating-rate profile on gfx950:
// BF16 MFMA GEMM (FP32 accumulate) with fused bias/ReLU epilogue, the
// elementwise kernels of training (ReLU backward, MSE gradient, bias
// gradient, Adam) and the column feature/target encoders.  No counterpart in
// the reference (the surrogate is new; it learns the mapping the exact
// kernels compute and is judged against them).
//
// GEMM: C[m][n] = sum_k A(m, k) B(k, n); A(m, k) = A[m*lda + k] (row-major) or
// A[k*lda + m] (transposed), B(k, n) = B[n*ldb + k] (K-contiguous per output
// column, the weight layout [out][in]) or B[k*ldb + n].  A 256-thread
// workgroup computes a 64 x 64 tile; each of its four waves a 32 x 32 block as
// 2 x 2 MFMA 16x16x16 tiles.  K advances in steps of 16 through LDS tiles
// stored [row][k], which absorbs the transposes at load time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/rrtmg_hip.h"
#include "mlp.hpp"

namespace rrtmg {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// raw feature f of a column from its loaded input value: T, ln q, O3 x 1e6,
// Ts, cos zenith (ln q in FP32: the feature is rounded to FP32 and then to
// BF16 after normalisation, and the FP32 logarithm is a single instruction
// sequence where the FP64 one is a long polynomial)
__device__ __forceinline__ float mlp_feature(int f, int nlay, double v)
{
    if (f < nlay) return (float)v;
    if (f < 2 * nlay) return logf((float)v);
    if (f < 3 * nlay) return (float)(v * 1.e6);
    return (float)v;
}

constexpr int GT = 64;      // output tile
constexpr int GK = 16;      // K step
constexpr int LDK = GK + 4; // LDS row pitch (bf16 elements)

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
                                                        const uint16_t *__restrict__ B, int ldb, GemmEpilogue ep)
{
    __shared__ uint16_t As[GT * LDK];
    __shared__ uint16_t Bs[GT * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += GK) {
        for (int e = tid; e < GT * GK; e += 256) {
            int m, k;
            if (TA) { m = e % GT; k = e / GT; } else { m = e / GK; k = e % GK; }
            const int gm = m0 + m, gk = k0 + k;
            uint16_t v = 0;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[m * LDK + k] = v;
            int n;
            if (TB) { n = e % GT; k = e / GT; } else { n = e / GK; k = e % GK; }
            const int gn = n0 + n;
            const int gk2 = k0 + k;
            uint16_t u = 0;
            if (gn < N && gk2 < K) u = TB ? B[(size_t)gk2 * ldb + gn] : B[(size_t)gn * ldb + gk2];
            Bs[n * LDK + k] = u;
        }
        __syncthreads();
        const int kq = 4 * (lane >> 4);
        for (int i = 0; i < 2; ++i) {
            const uint16_t *ap = As + (wm + 16 * i + (lane & 15)) * LDK + kq;
            bf16x4 a = {(short)ap[0], (short)ap[1], (short)ap[2], (short)ap[3]};
            for (int j = 0; j < 2; ++j) {
                const uint16_t *bp = Bs + (wn + 16 * j + (lane & 15)) * LDK + kq;
                bf16x4 b = {(short)bp[0], (short)bp[1], (short)bp[2], (short)bp[3]};
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // C lane layout of the 16x16 MFMA tile: row 4*(lane/16) + r, column lane%16
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= M || n >= N) continue;
                float v = acc[i][j][r];
                if (ep.bias) v += ep.bias[n];
                const size_t o = (size_t)m * ep.ldc + n;
                if (ep.out_pre) ep.out_pre[o] = v;
                if (ep.relu) v = v > 0.f ? v : 0.f;
                if (ep.out_f32) ep.out_f32[o] = ep.accumulate ? ep.out_f32[o] + v : v;
                if (ep.out_bf16) ep.out_bf16[o] = f2bf(v);
            }
}

// ---------------------------------------------------------------------------
// C = A B^T with both operands K-contiguous (A [M][K] activations, B [N][K]
// weights): the network's forward GEMM and dX = dY W^T's layout after
// transposition is not needed.  Workgroup tile 128 x 128, four waves of
// 64 x 64 (4 x 4 MFMA 16x16x16 tiles, 64 accumulator registers per lane);
// K in steps of 32 staged through LDS with 16-byte global loads, the next
// step's loads issued before the current step's MFMAs.
constexpr int NB = 128;       // output tile
constexpr int NK = 32;        // K step
constexpr int NLD = NK + 8;   // LDS row pitch (bf16): 80 bytes, rows on different banks

typedef __attribute__((__vector_size__(4 * sizeof(unsigned)))) unsigned u32x4;

__global__ __launch_bounds__(256) void gemm_bf16_nn_kernel(int M, int N, int K, const uint16_t *__restrict__ A,
                                                           int lda, const uint16_t *__restrict__ B, int ldb,
                                                           GemmEpilogue ep)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[NB * NLD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[NB * NLD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * NB, n0 = blockIdx.x * NB;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x4 acc[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    // each thread moves 2 chunks of 8 bf16 per operand and step: chunk c =
    // tid + 256 q -> row c / 4, k offset 8 (c % 4)
    u32x4 ra[2], rb[2];
    // 8 bf16 of row `row` of a K-contiguous operand from k, zero beyond the
    // matrix (the vector load when the whole chunk is inside)
    auto chunk = [&](const uint16_t *__restrict__ X, int ldx_, int rows, int row, int k) -> u32x4 {
        if (row < rows && k + 8 <= K) return *(const u32x4 *)(X + (size_t)row * ldx_ + k);
        u32x4 v = (u32x4){0u, 0u, 0u, 0u};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const unsigned h = (row < rows && k + e < K) ? X[(size_t)row * ldx_ + k + e] : 0u;
            v[e >> 1] |= h << (16 * (e & 1));
        }
        return v;
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int c = tid + 256 * q, row = c >> 2, kc = k0 + 8 * (c & 3);
            ra[q] = chunk(A, lda, M, m0 + row, kc);
            rb[q] = chunk(B, ldb, N, n0 + row, kc);
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += NK) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int c = tid + 256 * q, row = c >> 2, kk = 8 * (c & 3);
            *(u32x4 *)(As + row * NLD + kk) = ra[q];
            *(u32x4 *)(Bs + row * NLD + kk) = rb[q];
        }
        __syncthreads();
        if (k0 + NK < K) load(k0 + NK);  // next step's loads in flight during the MFMAs
#pragma unroll
        for (int kq = 0; kq < NK; kq += 16) {
            const int ko = kq + 4 * (lane >> 4);
            bf16x4 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = *(const bf16x4 *)(As + (wm + 16 * i + (lane & 15)) * NLD + ko);
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = *(const bf16x4 *)(Bs + (wn + 16 * j + (lane & 15)) * NLD + ko);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // C lane layout of the 16x16 MFMA tile: row 4*(lane/16) + r, column lane%16
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= M || n >= N) continue;
                float v = acc[i][j][r];
                if (ep.bias) v += ep.bias[n];
                const size_t o = (size_t)m * ep.ldc + n;
                if (ep.out_pre) ep.out_pre[o] = v;
                if (ep.relu) v = v > 0.f ? v : 0.f;
                if (ep.out_f32) ep.out_f32[o] = ep.accumulate ? ep.out_f32[o] + v : v;
                if (ep.out_bf16) ep.out_bf16[o] = f2bf(v);
            }
}

// ---------------------------------------------------------------------------
// C = A^T B with both operands stored [k][*] (A(m, k) = A[k*lda + m], B(k, n)
// = B[k*ldb + n]): the weight gradient dW = dY^T X, whose K is the batch and
// whose 64 x 64 output tiles are too few to fill the chip (16 for 256 x 256).
// K is split over grid.z; each workgroup accumulates its K range of one tile
// (16-byte loads along m and n, K in steps of 32 through LDS) and writes the
// FP32 partial to part[z][m][n]; splitk_reduce_kernel sums the partials in z
// order (deterministic) and applies the epilogue.
constexpr int TK = 32;        // K step
constexpr int TLD = TK + 8;   // LDS row pitch (bf16)

__global__ __launch_bounds__(256) void gemm_bf16_tt_splitk_kernel(int M, int N, int K, int kchunk,
                                                                  const uint16_t *__restrict__ A, int lda,
                                                                  const uint16_t *__restrict__ B, int ldb,
                                                                  float *__restrict__ part)
{
    __shared__ uint16_t As[GT * TLD];
    __shared__ uint16_t Bs[GT * TLD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    // loader: thread t moves 8 consecutive m (or n) of row k = t / 8 of the step
    const int lk = tid >> 3, lc = (tid & 7) * 8;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        const int gk = k0 + lk;
        u32x4 va = {0u, 0u, 0u, 0u}, vb = {0u, 0u, 0u, 0u};
        if (gk < kend) {
            if (m0 + lc < M) va = *(const u32x4 *)(A + (size_t)gk * lda + m0 + lc);
            if (n0 + lc < N) vb = *(const u32x4 *)(B + (size_t)gk * ldb + n0 + lc);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            As[(lc + 2 * q) * TLD + lk] = (uint16_t)(va[q] & 0xffffu);
            As[(lc + 2 * q + 1) * TLD + lk] = (uint16_t)(va[q] >> 16);
            Bs[(lc + 2 * q) * TLD + lk] = (uint16_t)(vb[q] & 0xffffu);
            Bs[(lc + 2 * q + 1) * TLD + lk] = (uint16_t)(vb[q] >> 16);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk += 16) {
            const int kq = kk + 4 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const uint16_t *ap = As + (wm + 16 * i + (lane & 15)) * TLD + kq;
                bf16x4 a = {(short)ap[0], (short)ap[1], (short)ap[2], (short)ap[3]};
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const uint16_t *bp = Bs + (wn + 16 * j + (lane & 15)) * TLD + kq;
                    bf16x4 b = {(short)bp[0], (short)bp[1], (short)bp[2], (short)bp[3]};
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float *pz = part + (size_t)blockIdx.z * M * N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m < M && n < N) pz[(size_t)m * N + n] = acc[i][j][r];
            }
}

__global__ __launch_bounds__(256) void splitk_reduce_kernel(int M, int N, int ks, const float *__restrict__ part,
                                                            GemmEpilogue ep)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)M * N) return;
    const int m = (int)(i / N), n = (int)(i - (size_t)m * N);
    float v = 0.f;
    for (int z = 0; z < ks; ++z) v += part[(size_t)z * M * N + i];
    if (ep.bias) v += ep.bias[n];
    const size_t o = (size_t)m * ep.ldc + n;
    if (ep.out_pre) ep.out_pre[o] = v;
    if (ep.relu) v = v > 0.f ? v : 0.f;
    if (ep.out_f32) ep.out_f32[o] = ep.accumulate ? ep.out_f32[o] + v : v;
    if (ep.out_bf16) ep.out_bf16[o] = f2bf(v);
}

int gemm_splitk_parts(int M, int N, int K, int transA, int transB, int lda, int ldb, const void *A, const void *B)
{
    if (!transA || !transB || M % 8 || N % 8 || lda % 8 || ldb % 8 || ((uintptr_t)A % 16) || ((uintptr_t)B % 16))
        return 0;
    const int tiles = ((M + GT - 1) / GT) * ((N + GT - 1) / GT);
    int ks = 256 / tiles;                      // about one workgroup per CU
    ks = std::min(ks, (K + 2 * TK - 1) / (2 * TK));  // at least two K steps per split
    return ks >= 2 ? ks : 0;
}

hipError_t launch_gemm_bf16(int M, int N, int K, const uint16_t *A, int lda, int transA, const uint16_t *B, int ldb,
                            int transB, const GemmEpilogue &ep, hipStream_t s, float *part, int ks)
{
    if (ks >= 2 && part) {
        const int kchunk = ((K + ks - 1) / ks + TK - 1) / TK * TK;
        const int kz = (K + kchunk - 1) / kchunk;
        dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, kz), block(256);
        hipLaunchKernelGGL(gemm_bf16_tt_splitk_kernel, grid, block, 0, s, M, N, K, kchunk, A, lda, B, ldb, part);
        const int nb = (int)(((size_t)M * N + 255) / 256);
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3(nb), dim3(256), 0, s, M, N, kz, part, ep);
        return hipGetLastError();
    }
    // both operands K-contiguous with 16-byte aligned rows: the tiled kernel
    if (!transA && !transB && lda % 8 == 0 && ldb % 8 == 0 && ((uintptr_t)A % 16) == 0 && ((uintptr_t)B % 16) == 0) {
        dim3 grid((N + NB - 1) / NB, (M + NB - 1) / NB), block(256);
        hipLaunchKernelGGL(gemm_bf16_nn_kernel, grid, block, 0, s, M, N, K, A, lda, B, ldb, ep);
        return hipGetLastError();
    }
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT), block(256);
    if (transA) {
        if (transB) hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, block, 0, s, M, N, K, A, lda, B, ldb, ep);
        else hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, block, 0, s, M, N, K, A, lda, B, ldb, ep);
    } else {
        if (transB) hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, block, 0, s, M, N, K, A, lda, B, ldb, ep);
        else hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, block, 0, s, M, N, K, A, lda, B, ldb, ep);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// elementwise kernels of training

// dZ = dY * (pre > 0) (ReLU backward), BF16 and optionally FP32
__global__ void relu_backward_kernel(size_t n, const float *__restrict__ dy, const float *__restrict__ pre,
                                     uint16_t *__restrict__ dz, float *__restrict__ dz32)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = pre[i] > 0.f ? dy[i] : 0.f;
        dz[i] = f2bf(v);
        if (dz32) dz32[i] = v;
    }
}

// per-output-column sum over rows (bias gradient): one thread per column
__global__ __launch_bounds__(1024) void col_sum_kernel(int M, int N, const float *__restrict__ x, int ldx,
                                                       float *__restrict__ out)
{
    // 64 columns per block, 16 row groups: thread (c, r) sums rows r, r + 16, ...
    // of column c; the 16 partial sums are added in a fixed order (deterministic)
    __shared__ float part[16][64];
    const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    float s = 0.f;
    // (unrolled: eight loads in flight, the additions in the same order)
    if (n < N) {
#pragma unroll 8
        for (int m = r; m < M; m += 16) s += x[(size_t)m * ldx + n];
    }
    part[r][c] = s;
    __syncthreads();
    if (r == 0 && n < N) {
        float t = 0.f;
        for (int j = 0; j < 16; ++j) t += part[j][c];
        out[n] = t;
    }
}

// mean squared error over M x N (row pitch ld) and its gradient 2 (y - t) / (M N)
// into dy (FP32) and dy_bf16; partial sums of the loss per block in loss[blockIdx]
__global__ void mse_kernel(int M, int N, const float *__restrict__ y, int ldy, const float *__restrict__ t, int ldt,
                           float *__restrict__ dy, uint16_t *__restrict__ dy_bf16, int ldd, float *__restrict__ loss)
{
    __shared__ float red[256];
    const size_t total = (size_t)M * N;
    const float scale = 2.f / (float)total;
    float s = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t m = i / N, n = i % N;
        const float d = y[m * ldy + n] - t[m * ldt + n];
        s += d * d;
        const float g = scale * d;
        if (dy) dy[m * ldd + n] = g;
        if (dy_bf16) dy_bf16[m * ldd + n] = f2bf(g);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[blockIdx.x] = red[0];
}

// Adam on FP32 master weights; writes the BF16 copy (padding stays zero)
__global__ void adam_kernel(size_t n, float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, uint16_t *__restrict__ wbf, AdamParams p)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i] + p.weight_decay * w[i];
        const float mi = p.beta1 * m[i] + (1.f - p.beta1) * gi;
        const float vi = p.beta2 * v[i] + (1.f - p.beta2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi * p.bc1, vh = vi * p.bc2;
        const float wi = w[i] - p.lr * mh / (sqrtf(vh) + p.eps);
        w[i] = wi;
        if (wbf) wbf[i] = f2bf(wi);
    }
}

// ---------------------------------------------------------------------------
// features of a column: T (nlay), ln q (nlay), O3 (nlay), Ts, cos zenith,
// normalised with per-feature (mean, 1/std), BF16, row-major [ncol][ldx]
// with zero padding; envelope flag = some feature outside [lo, hi]
// The three layout kernels below move between the library's level-major
// column arrays ([level][column], FP64) and the network's row-major
// [column][feature] arrays through an LDS tile of 64 columns, so that both
// the reads and the writes are coalesced.  Block: 256 threads = 64 columns x 4.

// features of 64 columns: T, ln q, O3 x 1e6 (nlay each), Ts, cos zenith;
// x_bf16 [ncol][ldx] (normalised when mean/istd are given), xraw (FP32, may be
// NULL), flag = 1 where some raw feature lies outside [lo, hi]
__global__ __launch_bounds__(256) void mlp_features_kernel(int ncol, int nlay, const double *__restrict__ tlay,
                                                           const double *__restrict__ h2o,
                                                           const double *__restrict__ o3,
                                                           const double *__restrict__ tsfc,
                                                           const double *__restrict__ coszen,
                                                           const float *__restrict__ mean,
                                                           const float *__restrict__ istd,
                                                           const float *__restrict__ lo, const float *__restrict__ hi,
                                                           int ldx, uint16_t *__restrict__ x, float *__restrict__ xraw,
                                                           int *__restrict__ flag)
{
    extern __shared__ float tile[];  // [64][ldx + 1], then lo, hi, mean, istd [ldx] each
    __shared__ int fl[64];
    const int ld = ldx + 1;
    const int col0 = blockIdx.x * 64;
    const int t = threadIdx.x, c = t & 63, q = t >> 6;
    const int col = col0 + c;
    const int nf = 3 * nlay + 2;
    // the per-feature bounds and scalings in LDS: read per feature below,
    // from global memory each would be a serial round trip
    float *s_lo = tile + 64 * ld, *s_hi = s_lo + ldx, *s_mean = s_hi + ldx, *s_istd = s_mean + ldx;
    for (int f = t; f < ldx; f += 256) {
        s_lo[f] = lo ? lo[f] : 0.f;
        s_hi[f] = lo ? hi[f] : 0.f;
        s_mean[f] = mean ? mean[f] : 0.f;
        s_istd[f] = mean ? istd[f] : 1.f;
    }
    if (t < 64) fl[t] = 0;
    __syncthreads();
    // raw inputs of 8 features are loaded before any is transformed, so each
    // wave keeps 8 loads in flight instead of one
    // (branch-free: source array and offset are selected, then one load, so
    // the 8 loads of a batch are not separated by waits at branch joins)
    auto raw = [&](int f) -> double {
        const int part = f < nlay ? 0 : f < 2 * nlay ? 1 : f < 3 * nlay ? 2 : f == 3 * nlay ? 3 : 4;
        const double *base = part == 0 ? tlay : part == 1 ? h2o : part == 2 ? o3 : part == 3 ? tsfc : coszen;
        const size_t off = part < 3 ? (size_t)(f - part * nlay) * ncol + col : (size_t)col;
        const bool ok = f < nf && col < ncol;
        return base[ok ? off : 0];
    };
    int out = 0;
    for (int f0 = q; f0 < ldx; f0 += 32) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = raw(f0 + 4 * u);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int f = f0 + 4 * u;
            if (f >= ldx) break;
            float r = 0.f;
            if (f < nf && col < ncol) {
                r = mlp_feature(f, nlay, v[u]);
                if (lo && (r < s_lo[f] || r > s_hi[f])) out = 1;
            }
            tile[c * ld + f] = r;
        }
    }
    if (out) atomicOr(&fl[c], 1);
    __syncthreads();
    // rows out, two BF16 features per 32-bit word (ldx is even)
    const int wpr = ldx / 2;
    uint32_t *x32 = (uint32_t *)x;
    for (int w = t; w < 64 * wpr; w += 256) {
        const int cc = w / wpr, j = w - cc * wpr, f0 = 2 * j;
        if (col0 + cc >= ncol) continue;
        const float r0 = tile[cc * ld + f0], r1 = tile[cc * ld + f0 + 1];
        const float z0 = f0 < nf ? (r0 - s_mean[f0]) * s_istd[f0] : 0.f;
        const float z1 = f0 + 1 < nf ? (r1 - s_mean[f0 + 1]) * s_istd[f0 + 1] : 0.f;
        x32[(size_t)(col0 + cc) * wpr + j] = (uint32_t)f2bf(z0) | ((uint32_t)f2bf(z1) << 16);
        if (xraw) {
            xraw[(size_t)(col0 + cc) * ldx + f0] = r0;
            xraw[(size_t)(col0 + cc) * ldx + f0 + 1] = r1;
        }
    }
    if (flag && t < 64 && col < ncol) flag[col] = fl[t];
}

// training targets: (hr_lw, hr_sw) of a column normalised per output, FP32 [ncol][ldt]
__global__ __launch_bounds__(256) void mlp_targets_kernel(int ncol, int nlay, const double *__restrict__ hr_lw,
                                                          const double *__restrict__ hr_sw,
                                                          const float *__restrict__ mean,
                                                          const float *__restrict__ istd, int ldt,
                                                          float *__restrict__ t)
{
    extern __shared__ float tile[];  // [64][ldt + 1], then mean, istd [ldt]
    const int ld = ldt + 1;
    const int col0 = blockIdx.x * 64;
    const int th = threadIdx.x, c = th & 63, q = th >> 6;
    const int col = col0 + c;
    float *s_mean = tile + 64 * ld, *s_istd = s_mean + ldt;
    for (int f = th; f < ldt; f += 256) {
        s_mean[f] = mean ? mean[f] : 0.f;
        s_istd[f] = mean ? istd[f] : 1.f;
    }
    __syncthreads();
    for (int f0 = q; f0 < ldt; f0 += 32) {
        double h[8];  // 8 loads in flight per wave
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int f = f0 + 4 * u;
            h[u] = f < 2 * nlay && col < ncol
                       ? (f < nlay ? hr_lw[(size_t)f * ncol + col] : hr_sw[(size_t)(f - nlay) * ncol + col])
                       : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int f = f0 + 4 * u;
            if (f >= ldt) break;
            tile[c * ld + f] = f < 2 * nlay ? ((float)h[u] - s_mean[f]) * s_istd[f] : 0.f;
        }
    }
    __syncthreads();
    for (int w = th; w < 64 * ldt; w += 256) {
        const int cc = w / ldt, f = w - cc * ldt;
        if (col0 + cc < ncol) t[(size_t)(col0 + cc) * ldt + f] = tile[cc * ld + f];
    }
}

// network outputs back to heating rates (level-major FP64, the library's layout)
__global__ __launch_bounds__(256) void mlp_decode_kernel(int ncol, int nlay, const float *__restrict__ y, int ldy,
                                                         const float *__restrict__ mean,
                                                         const float *__restrict__ std_, double *__restrict__ hr_lw,
                                                         double *__restrict__ hr_sw)
{
    extern __shared__ float tile[];  // [64][2 nlay + 1], then mean, std [2 nlay]
    const int nout = 2 * nlay, ld = nout + 1;
    const int col0 = blockIdx.x * 64;
    const int t = threadIdx.x, c = t & 63, q = t >> 6;
    float *s_mean = tile + 64 * ld, *s_std = s_mean + nout;
    for (int f = t; f < nout; f += 256) {
        s_mean[f] = mean[f];
        s_std[f] = std_[f];
    }
    for (int w0 = t; w0 < 64 * nout; w0 += 256 * 8) {
        float v[8];  // 8 loads in flight per wave
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int w = w0 + 256 * u, cc = w / nout, f = w - cc * nout;
            v[u] = w < 64 * nout && col0 + cc < ncol ? y[(size_t)(col0 + cc) * ldy + f] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int w = w0 + 256 * u, cc = w / nout, f = w - cc * nout;
            if (w < 64 * nout) tile[cc * ld + f] = v[u];
        }
    }
    __syncthreads();
    const int col = col0 + c;
    if (col >= ncol) return;
    for (int f = q; f < nout; f += 4) {
        const double h = (double)(tile[c * ld + f] * s_std[f] + s_mean[f]);
        if (f < nlay) hr_lw[(size_t)f * ncol + col] = h;
        else hr_sw[(size_t)(f - nlay) * ncol + col] = h;
    }
}

hipError_t launch_relu_backward(size_t n, const float *dy, const float *pre, uint16_t *dz, float *dz32, hipStream_t s)
{
    hipLaunchKernelGGL(relu_backward_kernel, dim3(1024), dim3(256), 0, s, n, dy, pre, dz, dz32);
    return hipGetLastError();
}
hipError_t launch_col_sum(int M, int N, const float *x, int ldx, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(col_sum_kernel, dim3((N + 63) / 64), dim3(1024), 0, s, M, N, x, ldx, out);
    return hipGetLastError();
}
hipError_t launch_mse(int M, int N, const float *y, int ldy, const float *t, int ldt, float *dy, uint16_t *dy_bf16,
                      int ldd, float *loss, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(mse_kernel, dim3(nblocks), dim3(256), 0, s, M, N, y, ldy, t, ldt, dy, dy_bf16, ldd, loss);
    return hipGetLastError();
}
hipError_t launch_adam(size_t n, float *w, float *m, float *v, const float *g, uint16_t *wbf, const AdamParams &p,
                       hipStream_t s)
{
    hipLaunchKernelGGL(adam_kernel, dim3(512), dim3(256), 0, s, n, w, m, v, g, wbf, p);
    return hipGetLastError();
}
hipError_t launch_mlp_features(int ncol, int nlay, const double *tlay, const double *h2o, const double *o3,
                               const double *tsfc, const double *coszen, const float *mean, const float *istd,
                               const float *lo, const float *hi, int ldx, uint16_t *x, float *xraw, int *flag,
                               hipStream_t s)
{
    if (ldx % 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_features_kernel, dim3((ncol + 63) / 64), dim3(256), sizeof(float) * (64 * (ldx + 1) + 4 * ldx), s, ncol,
                       nlay, tlay, h2o, o3, tsfc, coszen, mean, istd, lo, hi, ldx, x, xraw, flag);
    return hipGetLastError();
}
hipError_t launch_mlp_targets(int ncol, int nlay, const double *hr_lw, const double *hr_sw, const float *mean,
                              const float *istd, int ldt, float *t, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_targets_kernel, dim3((ncol + 63) / 64), dim3(256), sizeof(float) * (64 * (ldt + 1) + 2 * ldt), s, ncol,
                       nlay, hr_lw, hr_sw, mean, istd, ldt, t);
    return hipGetLastError();
}
hipError_t launch_mlp_decode(int ncol, int nlay, const float *y, int ldy, const float *mean, const float *std_,
                             double *hr_lw, double *hr_sw, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_decode_kernel, dim3((ncol + 63) / 64), dim3(256), sizeof(float) * (64 * (2 * nlay + 1) + 4 * nlay), s,
                       ncol, nlay, y, ldy, mean, std_, hr_lw, hr_sw);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Inference of the whole network in one kernel: a workgroup of four waves
// takes 64 columns, forms their features in LDS (BF16, normalised, with the
// envelope test), runs every layer with its activations kept in LDS and the
// weights read from L2, and writes the decoded heating rates.  Nothing but
// the inputs, the weights and the outputs touches memory.
//
// MFMA: gfx950's v_mfma_f32_16x16x32_bf16 (8 BF16 per lane and operand): lane
// l holds row (A) or column (B) l % 16 and k = 8 (l / 16) .. 8 (l / 16) + 7;
// the 16 x 16 FP32 result holds column l % 16, rows 4 (l / 16) + 0..3.  Each
// wave computes all four 16-row tiles of the 64 rows for the output column
// tiles wave, wave + 4, ... (at most four: widths up to 256), so the weight
// fragments of a K step are read once per workgroup and the activation
// fragments once per wave.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));

constexpr int FR = 64;            // columns (rows of the GEMMs) per workgroup
constexpr int FMAXW = 256;        // largest layer width
constexpr int FLDA = FMAXW + 8;   // activation row pitch (bf16): 528 bytes, 16 lanes on distinct banks
constexpr int FKS = FMAXW / 32;   // K steps of the widest layer input

struct FusedNet {
    int nlay, nin, nout, ldx, nlayers;
    int width[RRTMG_HIP_MLP_MAX_LAYERS + 1];  // width[0] = ldx, width[l + 1] = layer l's outputs
    const uint16_t *w[RRTMG_HIP_MLP_MAX_LAYERS];
    const float *b[RRTMG_HIP_MLP_MAX_LAYERS];
    const float *in_mean, *in_istd, *in_lo, *in_hi, *out_mean, *out_std;
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void mlp_infer_fused_kernel(int ncol, const double *__restrict__ tlay,
                                                              const double *__restrict__ h2o,
                                                              const double *__restrict__ o3,
                                                              const double *__restrict__ tsfc,
                                                              const double *__restrict__ coszen, FusedNet net,
                                                              double *__restrict__ hr_lw, double *__restrict__ hr_sw,
                                                              int *__restrict__ flag)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char fsm[];
    uint16_t *const act0 = (uint16_t *)fsm;
    uint16_t *const act1 = act0 + FR * FLDA;
    float *s_mean = (float *)(fsm + 2 * FR * FLDA * sizeof(uint16_t));
    float *s_istd = s_mean + FMAXW, *s_lo = s_istd + FMAXW, *s_hi = s_lo + FMAXW;
    int *fl = (int *)(s_hi + FMAXW);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int col0 = blockIdx.x * FR;
    const int nlay = net.nlay, nf = net.nin, ldx = net.ldx;
    for (int f = t; f < ldx; f += 256) {
        s_mean[f] = net.in_mean[f];
        s_istd[f] = net.in_istd[f];
        s_lo[f] = net.in_lo ? net.in_lo[f] : 0.f;
        s_hi[f] = net.in_lo ? net.in_hi[f] : 0.f;
    }
    if (t < FR) fl[t] = 0;
    const int lr = lane & 15, lk = 8 * (lane >> 4);
    // weight fragments of one layer for this wave: column tile j = wave + 4 j'
    // (tiles beyond the layer read the last tile; their MFMAs are skipped),
    // K step st; every load unconditional (K steps beyond win re-read the
    // last one), so the compiler waits for them only at their first use
    bf16x8 bw[4][FKS];
    auto wload = [&](int l) {
        // net.w[l] holds the layer's weights in fragment order (pack kernel
        // below): [column tile][K step][lane][8], one contiguous 1 KB per
        // wave and load
        const int nst = net.width[l] / 32, nct = net.width[l + 1] / 16;
        const bf16x8 *__restrict__ P = (const bf16x8 *)net.w[l];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ct = min(wave + 4 * j, nct - 1);
#pragma unroll
            for (int st = 0; st < FKS; ++st) bw[j][st] = P[((size_t)ct * nst + min(st, nst - 1)) * 64 + lane];
        }
    };
    __syncthreads();
    // features (as mlp_features_kernel forms them): thread (c, q) takes
    // features q, q + 4, ...; eight loads in flight before any transform
    {
        const int c = t & 63, q = t >> 6, col = col0 + c;
        const bool live = col < ncol;
        auto raw = [&](int f) -> double {
            const int part = f < nlay ? 0 : f < 2 * nlay ? 1 : f < 3 * nlay ? 2 : f == 3 * nlay ? 3 : 4;
            const double *base = part == 0 ? tlay : part == 1 ? h2o : part == 2 ? o3 : part == 3 ? tsfc : coszen;
            const size_t off = part < 3 ? (size_t)(f - part * nlay) * ncol + col : (size_t)col;
            return base[f < nf && live ? off : 0];
        };
        int out = 0;
        // a thread's features q, q + 4, ..., all 32 loads in flight at once
        // (ldx <= 128; wider feature sets in further batches)
        for (int f0 = q; f0 < ldx; f0 += 128) {
            double v[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) v[u] = raw(f0 + 4 * u);
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int f = f0 + 4 * u;
                if (f >= ldx) continue;
                float z = 0.f;
                if (f < nf && live) {
                    const float r = mlp_feature(f, nlay, v[u]);
                    if (net.in_lo && (r < s_lo[f] || r > s_hi[f])) out = 1;
                    z = (r - s_mean[f]) * s_istd[f];
                }
                act0[c * FLDA + f] = f2bf(z);
            }
        }
        if (out) atomicOr(&fl[c], 1);
    }
    wload(0);
    __syncthreads();
    int cur = 0;
    float *y = nullptr;  // the last layer's FP32 outputs, [FR][nout + 1], in the buffer it does not read
    const int ldy = net.nout + 1;
    for (int l = 0; l < net.nlayers; ++l) {
        const int win = net.width[l], wout = net.width[l + 1], nct = wout / 16;
        const bool last = l == net.nlayers - 1;
        const uint16_t *in = cur ? act1 : act0;
        f32x4v acc[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[j][r] = (f32x4v){0.f, 0.f, 0.f, 0.f};
        // the wave's weight fragments of the whole layer are in registers
        // (loaded before the previous layer's epilogue, or before the
        // features for layer 0): one L2 round trip per layer
#pragma unroll
        for (int st = 0; st < FKS; ++st) {
            if (st * 32 >= win) break;  // uniform
            bf16x8 a[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = *(const bf16x8 *)(in + (r * 16 + lr) * FLDA + st * 32 + lk);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (wave + 4 * j >= nct) break;  // uniform per wave
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[j][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], bw[j][st], acc[j][r], 0, 0, 0);
            }
        }
        __syncthreads();  // every wave is done reading `in` (the next buffer may alias it)
        uint16_t *outb = cur ? act0 : act1;
        if (last) y = (float *)outb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ct = wave + 4 * j;
            if (ct >= nct) break;
            const int n = ct * 16 + lr;
            const float bias = net.b[l][n];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int m = r * 16 + 4 * (lane >> 4) + e;
                    float v = acc[j][r][e] + bias;
                    if (last) {
                        y[m * ldy + n] = v;
                    } else {
                        v = v > 0.f ? v : 0.f;
                        outb[m * FLDA + n] = f2bf(v);
                    }
                }
        }
        if (!last) wload(l + 1);  // the next layer's weights, in flight across the barrier
        __syncthreads();
        cur = 1 - cur;
    }
    // decode: (hr_lw, hr_sw) level-major FP64; lanes over columns, coalesced
    {
        const int c = t & 63, q = t >> 6, col = col0 + c;
        if (col < ncol) {
            for (int f = q; f < net.nout; f += 4) {
                const double h = (double)(y[c * ldy + f] * net.out_std[f] + net.out_mean[f]);
                if (f < nlay) hr_lw[(size_t)f * ncol + col] = h;
                else hr_sw[(size_t)(f - nlay) * ncol + col] = h;
            }
            if (flag && q == 0) flag[col] = fl[c];
        }
    }
}

// W [wout][win] (row-major, K-contiguous) -> P [wout/16][win/32][64][8]:
// element e of lane l of (column tile ct, K step st) is
// W[ct*16 + l%16][st*32 + 8*(l/16) + e], the B operand of the fused kernel's
// v_mfma_f32_16x16x32_bf16 for that tile and step
__global__ __launch_bounds__(256) void mlp_pack_kernel(int wout, int win, const uint16_t *__restrict__ W,
                                                       uint16_t *__restrict__ P)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // one 8-element fragment row per thread
    const int nst = win / 32;
    if (i >= (size_t)(wout / 16) * nst * 64) return;
    const int lane = (int)(i % 64), st = (int)((i / 64) % nst), ct = (int)(i / 64 / nst);
    const uint16_t *src = W + (size_t)(ct * 16 + lane % 16) * win + st * 32 + 8 * (lane / 16);
    uint16_t *dst = P + i * 8;
#pragma unroll
    for (int e = 0; e < 8; ++e) dst[e] = src[e];
}

hipError_t launch_mlp_pack(int wout, int win, const uint16_t *W, uint16_t *P, hipStream_t s)
{
    const size_t n = (size_t)(wout / 16) * (win / 32) * 64;
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, wout, win, W, P);
    return hipGetLastError();
}

bool mlp_fused_fits(int ldx, int nlayers, const int *width, int nout)
{
    // every layer's input width a multiple of 32 (whole K steps), every
    // output width a multiple of 16 (whole column tiles), all at most 256
    if (nlayers < 1 || nlayers > RRTMG_HIP_MLP_MAX_LAYERS || ldx % 32 || ldx > FMAXW) return false;
    for (int l = 0; l < nlayers; ++l)
        if (width[l] < 16 || width[l] % (l + 1 < nlayers ? 32 : 16) || width[l] > FMAXW) return false;
    // the last layer's FP32 outputs, [FR][nout + 1], reuse one BF16 activation buffer
    return nout <= width[nlayers - 1] && (size_t)FR * (nout + 1) * sizeof(float) <= (size_t)FR * FLDA * 2;
}

size_t mlp_fused_lds_bytes() { return 2 * (size_t)FR * FLDA * sizeof(uint16_t) + 4 * FMAXW * sizeof(float) + FR * sizeof(int); }

hipError_t launch_mlp_infer_fused(int ncol, int nlay, const double *tlay, const double *h2o, const double *o3,
                                  const double *tsfc, const double *coszen, int ldx, int nlayers, const int *width,
                                  const void *const *w, const float *const *b, const float *in_mean,
                                  const float *in_istd, const float *in_lo, const float *in_hi,
                                  const float *out_mean, const float *out_std, double *hr_lw, double *hr_sw,
                                  int *flag, hipStream_t s)
{
    FusedNet n{};
    n.nlay = nlay;
    n.nin = 3 * nlay + 2;
    n.nout = 2 * nlay;
    n.ldx = ldx;
    n.nlayers = nlayers;
    n.width[0] = ldx;
    for (int l = 0; l < nlayers; ++l) {
        n.width[l + 1] = width[l];
        n.w[l] = (const uint16_t *)w[l];
        n.b[l] = b[l];
    }
    n.in_mean = in_mean;
    n.in_istd = in_istd;
    n.in_lo = in_lo;
    n.in_hi = in_hi;
    n.out_mean = out_mean;
    n.out_std = out_std;
    hipLaunchKernelGGL(mlp_infer_fused_kernel, dim3((ncol + FR - 1) / FR), dim3(256), mlp_fused_lds_bytes(), s, ncol,
                       tlay, h2o, o3, tsfc, coszen, n, hr_lw, hr_sw, flag);
    return hipGetLastError();
}

}  // namespace rrtmg
