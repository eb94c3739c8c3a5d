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

#include <cstdint>

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

hipError_t launch_gemm_bf16(int M, int N, int K, const uint16_t *A, int lda, int transA, const uint16_t *B, int ldb,
                            int transB, const GemmEpilogue &ep, hipStream_t s)
{
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
__global__ void col_sum_kernel(int M, int N, const float *__restrict__ x, int ldx, float *__restrict__ out)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += x[(size_t)m * ldx + n];
    out[n] = s;
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
__global__ void mlp_features_kernel(int ncol, int nlay, const double *__restrict__ tlay,
                                    const double *__restrict__ h2o, const double *__restrict__ o3,
                                    const double *__restrict__ tsfc, const double *__restrict__ coszen,
                                    const float *__restrict__ mean, const float *__restrict__ istd,
                                    const float *__restrict__ lo, const float *__restrict__ hi, int ldx,
                                    uint16_t *__restrict__ x, float *__restrict__ xraw, int *__restrict__ flag)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const int nf = 3 * nlay + 2;
    int out = 0;
    for (int f = 0; f < ldx; ++f) {
        float r = 0.f;
        if (f < nf) {
            double v;
            if (f < nlay) v = tlay[(size_t)f * ncol + col];
            else if (f < 2 * nlay) v = log(h2o[(size_t)(f - nlay) * ncol + col]);
            else if (f < 3 * nlay) v = o3[(size_t)(f - 2 * nlay) * ncol + col] * 1.e6;
            else if (f == 3 * nlay) v = tsfc[col];
            else v = coszen[col];
            r = (float)v;
            if (lo && (r < lo[f] || r > hi[f])) out = 1;
        }
        if (xraw) xraw[(size_t)col * ldx + f] = r;
        const float z = f < nf && mean ? (r - mean[f]) * istd[f] : (f < nf ? r : 0.f);
        x[(size_t)col * ldx + f] = f2bf(z);
    }
    if (flag) flag[col] = out;
}

// targets: (hr_lw, hr_sw) of a column normalised per output, FP32 [ncol][ldt]
__global__ void mlp_targets_kernel(int ncol, int nlay, const double *__restrict__ hr_lw,
                                   const double *__restrict__ hr_sw, const float *__restrict__ mean,
                                   const float *__restrict__ istd, int ldt, float *__restrict__ t)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    for (int f = 0; f < ldt; ++f) {
        float v = 0.f;
        if (f < 2 * nlay) {
            const double h = f < nlay ? hr_lw[(size_t)f * ncol + col] : hr_sw[(size_t)(f - nlay) * ncol + col];
            v = mean ? ((float)h - mean[f]) * istd[f] : (float)h;
        }
        t[(size_t)col * ldt + f] = v;
    }
}

// network outputs back to heating rates (level-major FP64, the library's layout)
__global__ void mlp_decode_kernel(int ncol, int nlay, const float *__restrict__ y, int ldy,
                                  const float *__restrict__ mean, const float *__restrict__ std_,
                                  double *__restrict__ hr_lw, double *__restrict__ hr_sw)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    for (int f = 0; f < 2 * nlay; ++f) {
        const double h = (double)(y[(size_t)col * ldy + f] * std_[f] + mean[f]);
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
    hipLaunchKernelGGL(col_sum_kernel, dim3((N + 63) / 64), dim3(64), 0, s, M, N, x, ldx, out);
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
    hipLaunchKernelGGL(mlp_features_kernel, dim3((ncol + 127) / 128), dim3(128), 0, s, ncol, nlay, tlay, h2o, o3, tsfc,
                       coszen, mean, istd, lo, hi, ldx, x, xraw, flag);
    return hipGetLastError();
}
hipError_t launch_mlp_targets(int ncol, int nlay, const double *hr_lw, const double *hr_sw, const float *mean,
                              const float *istd, int ldt, float *t, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_targets_kernel, dim3((ncol + 127) / 128), dim3(128), 0, s, ncol, nlay, hr_lw, hr_sw, mean,
                       istd, ldt, t);
    return hipGetLastError();
}
hipError_t launch_mlp_decode(int ncol, int nlay, const float *y, int ldy, const float *mean, const float *std_,
                             double *hr_lw, double *hr_sw, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_decode_kernel, dim3((ncol + 127) / 128), dim3(128), 0, s, ncol, nlay, y, ldy, mean, std_,
                       hr_lw, hr_sw);
    return hipGetLastError();
}

}  // namespace rrtmg
