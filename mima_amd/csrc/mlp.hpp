// Launch interfaces of the MLP surrogate kernels (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rrtmg {

// epilogue of one GEMM: v = acc (+ bias[n]); out_pre = v; v = relu(v) if relu;
// out_f32 = v (or += v with accumulate); out_bf16 = bf16(v); any output may be NULL
struct GemmEpilogue {
    const float *bias = nullptr;
    int relu = 0;
    float *out_pre = nullptr;
    float *out_f32 = nullptr;
    uint16_t *out_bf16 = nullptr;
    int ldc = 0;
    int accumulate = 0;
};

struct AdamParams {
    float lr, beta1, beta2, eps, weight_decay;
    float bc1, bc2;  // 1/(1 - beta1^t), 1/(1 - beta2^t)
};

// split-K factor of the A^T B kernel for this call (0: the direct kernels);
// the caller provides part, ks x M x N floats of workspace
int gemm_splitk_parts(int M, int N, int K, int transA, int transB, int lda, int ldb, const void *A, const void *B);
hipError_t launch_gemm_bf16(int M, int N, int K, const uint16_t *A, int lda, int transA, const uint16_t *B, int ldb,
                            int transB, const GemmEpilogue &ep, hipStream_t s, float *part = nullptr, int ks = 0);
hipError_t launch_relu_backward(size_t n, const float *dy, const float *pre, uint16_t *dz, float *dz32, hipStream_t s);
hipError_t launch_col_sum(int M, int N, const float *x, int ldx, float *out, hipStream_t s);
hipError_t launch_mse(int M, int N, const float *y, int ldy, const float *t, int ldt, float *dy, uint16_t *dy_bf16,
                      int ldd, float *loss, int nblocks, hipStream_t s);
hipError_t launch_adam(size_t n, float *w, float *m, float *v, const float *g, uint16_t *wbf, const AdamParams &p,
                       hipStream_t s);
hipError_t launch_mlp_features(int ncol, int nlay, const double *tlay, const double *h2o, const double *o3,
                               const double *tsfc, const double *coszen, const float *mean, const float *istd,
                               const float *lo, const float *hi, int ldx, uint16_t *x, float *xraw, int *flag,
                               hipStream_t s);
hipError_t launch_mlp_targets(int ncol, int nlay, const double *hr_lw, const double *hr_sw, const float *mean,
                              const float *istd, int ldt, float *t, hipStream_t s);
// the whole inference in one kernel (features, every layer, decode) for
// networks whose layer inputs are multiples of 32 and outputs multiples of 16,
// all at most 256; its weights w[l] in fragment order (launch_mlp_pack)
bool mlp_fused_fits(int ldx, int nlayers, const int *width, int nout);
hipError_t launch_mlp_pack(int wout, int win, const uint16_t *W, uint16_t *P, hipStream_t s);
hipError_t launch_mlp_infer_fused(int ncol, int nlay, const double *tlay, const double *h2o, const double *o3,
                                  const double *tsfc, const double *coszen, int ldx, int nlayers, const int *width,
                                  const void *const *w, const float *const *b, const float *in_mean,
                                  const float *in_istd, const float *in_lo, const float *in_hi,
                                  const float *out_mean, const float *out_std, double *hr_lw, double *hr_sw,
                                  int *flag, hipStream_t s);
hipError_t launch_mlp_decode(int ncol, int nlay, const float *y, int ldy, const float *mean, const float *std_,
                             double *hr_lw, double *hr_sw, hipStream_t s);

}  // namespace rrtmg
