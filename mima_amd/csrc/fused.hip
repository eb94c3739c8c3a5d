// The band phase of the one-pass forward+backward of both sides in one
// persistent kernel (FP32): LW and SW work items from one counter.  No
// counterpart in the reference.  The per-unit arithmetic is that of the LW
// LDS-table backward kernel and of the SW backward band kernel
// (lw_vjp_body.hpp, sw_vjp_body.hpp); what changes is the schedule: one launch
// whose workgroups move from LW items to SW items without a kernel boundary,
// instead of two kernels on two streams.  Measured at 131072 x 60 it is the
// slower schedule (88.1 against 81.2 ms per forward+backward step): the SW
// body needs over 168 VGPRs, so the shared kernel runs the LW items at two
// waves per SIMD where the LW kernel alone runs three.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "lw_vjp_body.hpp"
#include "rrtmg_device.hpp"
#include "sw_vjp_body.hpp"

namespace rrtmg {

size_t fused_lds_bytes(const LwLdsPlan &plan, int sw_nlay)
{
    return std::max((size_t)plan.lds_bytes, (size_t)FUSED_WGW * sw_vjp_lds_bytes(sw_nlay, sizeof(float)));
}

size_t fused_row_floats(int lw_nlay, int sw_nlay)
{
    return std::max(lw_vjp_lds_state_bytes(lw_nlay), sw_vjp_state_bytes(sw_nlay, sizeof(float))) / sizeof(float);
}

int fused_grid(const LwLdsPlan &plan, int nitems) { return std::min(plan.ncu, nitems); }

// each side's band body is a call of its own (lw_vjp_unit, sw_vjp_unit), so
// that the register allocation of one body does not weigh on the other's;
// lw_vjp_unit is also what the LW LDS-table backward kernel calls (bitwise
// the same LW results), the SW backward band kernel inlines its body

// 8 waves per workgroup at up to 256 VGPRs: two waves per SIMD
__global__ __launch_bounds__(64 * FUSED_WGW) __attribute__((amdgpu_waves_per_eu(2)))
void fused_band_kernel(FusedBand a)
{
    extern __shared__ unsigned char smem_raw[];
    float *lds = (float *)smem_raw;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    float *row = a.rows + ((size_t)blockIdx.x * FUSED_WGW + wave) * a.row_floats;
    const int lnlay = a.lin.nlay;
    const int ltiles = (a.lin.ncol + TILE_COLS - 1) / TILE_COLS;
    const size_t sw_wave_lds = sw_vjp_lds_bytes(a.sin.nlay, sizeof(float)) / sizeof(float);
    __shared__ int s_item;
    const SwUnitArgs<float> sa = sw_unit_args<float>(a.sin, a.sw, a.sexp_off, a.stab);
    const LwUnitArgs la = lw_unit_args(a.lin, a.lw, a.exp_off, a.tfn_off, a.plk_off, a.ltab, a.cst);
    int loaded = -1;  // the LW band whose k-tables LDS holds
    for (;;) {
        // (next_item synchronises the workgroup: every wave is done with the
        // previous item, and with what it kept in LDS, before the next starts)
        const int item = next_item(a.items, &s_item);
        if (item >= a.nitems) break;
        if (item < a.litems) {
            const int group = item / a.lnwaves, r = item - group * a.lnwaves;
            const int b = a.lunits[r].band;
            if (b != loaded) {
                const int t0 = a.plan.tab0[b], n = a.plan.tab0[b + 1] - t0;
#pragma unroll 8
                for (int i = threadIdx.x; i < n; i += 64 * FUSED_WGW) lds[i] = a.ltab[t0 + i];
                __syncthreads();
                loaded = b;
            }
            const int k = (r - a.plan.ubeg[b]) * FUSED_WGW + wave;
            const int tile = group * FUSED_WGW + k / a.plan.upt[b];
            if (tile >= ltiles) continue;
            const WaveUnit u = a.lunits[a.plan.ubeg[b] + k % a.plan.upt[b]];
            const LwBand B = a.lbands[b];
            float *s3 = row, *s1 = row + (size_t)2 * lnlay * 64;
#define LBODY(bb)                                                                                              \
    case bb - 1:                                                                                               \
        lw_vjp_unit<bb>(la, B, (lptr<const float>)lds, u, tile, lane, (gptr<float>)s1, (gptr<float>)s3);      \
        break;
            switch (b) {
                LBODY(1) LBODY(2) LBODY(3) LBODY(4) LBODY(5) LBODY(6) LBODY(7) LBODY(8)
                LBODY(9) LBODY(10) LBODY(11) LBODY(12) LBODY(13) LBODY(14) LBODY(15) LBODY(16)
            }
#undef LBODY
        } else {
            loaded = -1;  // the SW units' LDS overlays the k-tables
            const int blk = (item - a.litems) * FUSED_WGW + wave;
            if (blk >= a.snunits) continue;
            const int tile = blk / a.snwaves;
            const WaveUnit u = a.sunits[blk % a.snwaves];
            const SwBand B = a.sbands[u.band];
            float *smem = lds + (size_t)wave * sw_wave_lds;
#define SBODY(jb)                                                                                  \
    case jb - 16:                                                                                  \
        sw_vjp_unit<jb, float>(sa, B, u, tile, lane, (gptr<float>)row, (lptr<float>)smem);       \
        break;
            switch (u.band) { SW_BANDS(SBODY) }
#undef SBODY
        }
    }
}

hipError_t launch_fused_band(const FusedBand &a, hipStream_t s)
{
    hipLaunchKernelGGL(reset_items_kernel, dim3(1), dim3(1), 0, s, a.items);
    hipLaunchKernelGGL(fused_band_kernel, dim3(fused_grid(a.plan, a.nitems)), dim3(64 * FUSED_WGW),
                       fused_lds_bytes(a.plan, a.sin.nlay), s, a);
    return hipGetLastError();
}

}  // namespace rrtmg
