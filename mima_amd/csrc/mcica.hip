// McICA stochastic cloud subcolumns on gfx950: mcica_subcol_lw and
// mcica_subcol_sw with the KISS generator (irng = 0), the generator of both
// rrtmg_lw/gcm_model/src/mcica_subcol_gen_lw.f90 and
// rrtmg_sw/gcm_model/src/mcica_subcol_gen_sw.f90 (140 LW, 112 SW
// subcolumns).  The generator is sequential per column; its parts are
// linear, so each (column, subcolumn) thread jumps to its own place in the
// sequence (see below).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace rrtmg {

// One step of the KISS generator (mcica_subcol_gen_lw.f90:530-562): a
// congruential, a 3-shift register and two multiply-with-carry generators
// in 32-bit wrap-around arithmetic, their sum read as a signed integer and
// mapped to (0, 1) by a product and a sum rounded separately (no fused
// multiply-add), as the reference computes it.
__device__ __forceinline__ double kiss_step(uint32_t &s1, uint32_t &s2, uint32_t &s3, uint32_t &s4)
{
#pragma clang fp contract(off)
    s1 = 69069u * s1 + 1327217885u;
    uint32_t k = s2;
    k ^= k << 13;
    k ^= k >> 17;
    k ^= k << 5;
    s2 = k;
    s3 = 18000u * (s3 & 65535u) + (s3 >> 16);
    s4 = 30903u * (s4 & 65535u) + (s4 >> 16);
    const uint32_t sum = s1 + s2 + (s3 << 16) + s4;
    return (double)(int32_t)sum * 2.328306e-10 + 0.5;
}

// Jump-ahead of the KISS generator: each of its four parts is linear, so
// the state d steps on follows from the seed without the d steps between.
//   congruential s1 -> 69069 s1 + 1327217885 (mod 2^32): affine, its d-th
//     power A s1 + C by repeated squaring;
//   3-shift register: linear over GF(2), the d-th power of its 32 x 32 bit
//     matrix as a product of the squarings M^(2^k) (KISS_XS_JUMPS of them);
//   multiply-with-carry s -> a (s mod 2^16) + s / 2^16: with m = a 2^16 - 1,
//     2^16 s' = s (mod m), so s' = a s (mod m) and s_d = a^d s (mod m); from
//     the second step on the state lies in [0, m), except the fixed point m
//     reached from a nonzero multiple of m.
constexpr uint64_t MWC3_A = 18000u, MWC4_A = 30903u;
constexpr uint64_t MWC3_M = MWC3_A * 65536u - 1u, MWC4_M = MWC4_A * 65536u - 1u;

void kiss_xorshift_jumps(uint32_t *t)
{
    auto xs = [](uint32_t k) {
        k ^= k << 13;
        k ^= k >> 17;
        k ^= k << 5;
        return k;
    };
    auto apply = [](const uint32_t *m, uint32_t v) {
        uint32_t r = 0;
        for (int j = 0; j < 32; ++j)
            if ((v >> j) & 1u) r ^= m[j];
        return r;
    };
    for (int j = 0; j < 32; ++j) t[j] = xs(1u << j);
    for (int k = 1; k < KISS_XS_JUMPS; ++k)
        for (int j = 0; j < 32; ++j) t[k * 32 + j] = apply(t + (k - 1) * 32, t[(k - 1) * 32 + j]);
}

__device__ __forceinline__ uint64_t mulmod(uint64_t x, uint64_t y, uint64_t m) { return x * y % m; }

__device__ uint64_t powmod(uint64_t a, uint64_t e, uint64_t m)
{
    uint64_t r = 1 % m;
    a %= m;
    for (; e; e >>= 1) {
        if (e & 1u) r = mulmod(r, a, m);
        a = mulmod(a, a, m);
    }
    return r;
}

// the generator steps before subcolumn ig's first number: permuteseed, then
// nlay numbers per earlier subcolumn (one for maximum overlap)
__device__ __forceinline__ uint64_t mcica_skip(int permuteseed, int icld, int nlay, int ig)
{
    return (uint64_t)(permuteseed > 0 ? permuteseed : 0) + (uint64_t)ig * (uint64_t)(icld == 3 ? 1 : nlay);
}

// g-points of band b (0-based) of the side with nsub subcolumns (140 LW, 112 SW)
__device__ __forceinline__ int mcica_band_ng(int nsub, int b) { return nsub == 140 ? lw_ng(b + 1) : sw_ng(b + 16); }

// per subcolumn (one thread each): the jump constants of its skip and its band
__global__ __launch_bounds__(192) void mcica_jump_kernel(int nlay, int icld, int permuteseed, int nsub,
                                                         McicaJump *jump)
{
    const int ig = threadIdx.x;
    if (ig >= nsub) return;
    const uint64_t d = mcica_skip(permuteseed, icld, nlay, ig);
    uint32_t A = 1u, C = 0u, a = 69069u, c = 1327217885u;
    for (uint64_t e = d; e; e >>= 1) {
        if (e & 1u) {
            A = a * A;
            C = a * C + c;
        }
        c = a * c + c;
        a = a * a;
    }
    int b = 0, end = mcica_band_ng(nsub, 0);
    while (ig >= end) end += mcica_band_ng(nsub, ++b);
    McicaJump J;
    J.a1 = A;
    J.c1 = C;
    J.p3 = (uint32_t)powmod(MWC3_A, d, MWC3_M);
    J.p4 = (uint32_t)powmod(MWC4_A, d, MWC4_M);
    J.band = b;
    jump[ig] = J;
}

// s_d of a multiply-with-carry part from s_0 (p = a^d mod m)
__device__ __forceinline__ uint32_t mwc_jump(uint32_t s, uint64_t d, uint64_t p, uint64_t a, uint64_t m)
{
    if (d < 2) {
        for (uint64_t n = 0; n < d; ++n) s = (uint32_t)a * (s & 65535u) + (s >> 16);
        return s;
    }
    const uint32_t r = (uint32_t)mulmod(p, s % m, m);
    return r == 0 && s != 0 ? (uint32_t)m : r;
}

// mcica_subcol_lw / mcica_subcol_sw with the KISS generator
// (mcica_subcol_gen_lw.f90:48-152, generate_stochastic_clouds :158-520;
// mcica_subcol_gen_sw.f90:47-159, generate_stochastic_clouds_sw :163-536:
// the same generator over 140 or 112 subcolumns).  One thread per (column,
// subcolumn) in subcolumn-fastest order, so a wave's stores of a layer are
// contiguous: the thread seeds the generator from the fractional Pa of its
// column's four lowest layers, jumps over the numbers that precede its
// subcolumn (permuteseed, then nlay per earlier subcolumn, or one for
// maximum overlap), and draws its subcolumn's numbers in layer order: one
// per layer for random (1) and maximum-random (2) overlap, the latter
// carried up through cloudy layers; one for all layers for maximum overlap
// (3).  Cloudy where the number >= 1 - fraction, with the band's optics
// and the layer's water paths; a clear SW subcolumn has depth 0, albedo 1,
// asymmetry 0, forward fraction 0.  A column whose second layer has the
// higher pressure (the reference stops, mcica_subcol_gen_lw.f90:327-329) is
// counted in *bad (when given) and left unwritten.  NULL water paths or
// radii read as zero; NULL outputs other than cldfmcl and taucmcl are
// skipped.  Contraction is off: the reference rounds every product and sum
// on its own, and a fused play * 100 - int(play * 100) would keep bits it
// rounds away and change the seed.
// The Mersenne Twister of mcica_random_numbers.f90 (MT19937, Matsumoto and
// Nishimura 1998): one stream per generator call, seeded with permuteseed
// (mcica_subcol_gen_lw.f90:338-339; the scalar initialisation, :169-185)
// and drawn in the order subcolumn, column, layer (:360-367, :385-392,
// :418-425).  The stream is serial, so one workgroup produces it: per block
// of 624 words, the twist (:134-152) in four dependent phases over the
// state in LDS, then the tempered words (:154-165) are stored.
__device__ __forceinline__ uint32_t mt_twist(uint32_t u, uint32_t v)
{
    const uint32_t m = (u & 0x80000000u) | (v & 0x7fffffffu);
    return (m >> 1) ^ ((v & 1u) ? 0x9908b0dfu : 0u);
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    return y ^ (y >> 18);
}

__global__ __launch_bounds__(640) void mt_stream_kernel(uint32_t seed, size_t count, uint32_t *__restrict__ out)
{
    __shared__ uint32_t st[624];
    const int k = (int)threadIdx.x;
    if (k == 0) {
        uint32_t x = seed;
        st[0] = x;
        for (int i = 1; i < 624; ++i) {
            x = 1812433253u * (x ^ (x >> 30)) + (uint32_t)i;
            st[i] = x;
        }
    }
    __syncthreads();
    for (size_t base = 0; base < count; base += 624) {
        // the first draw twists first (the state starts with its element
        // counter at the block size)
        uint32_t t = 0, far = 0, last = 0;
        if (k < 623) t = mt_twist(st[k], st[k + 1]);
        if (k < 227) far = st[k + 397];
        if (k == 623) last = st[623];
        __syncthreads();
        if (k < 227) st[k] = far ^ t;
        __syncthreads();
        if (k >= 227 && k < 454) st[k] = st[k - 227] ^ t;
        __syncthreads();
        if (k >= 454 && k < 623) st[k] = st[k - 227] ^ t;
        __syncthreads();
        if (k == 623) st[623] = st[396] ^ mt_twist(last, st[0]);
        __syncthreads();
        if (k < 624 && base + k < count) out[base + k] = mt_temper(st[k]);
    }
}

// one (column, subcolumn) of the generator with the Mersenne Twister's
// numbers: getRandomReal (genrand_real1, mcica_random_numbers.f90:276-295)
// of the stream word, then the overlap as with KISS (no seeds, so no
// pressure-order condition)
__device__ __forceinline__ void mcica_subcol_mt(const McicaSubcolArgs &a, int col, int ig)
{
#pragma clang fp contract(off)
    const int ncol = a.ncol, nlay = a.nlay, icld = a.icld, nsub = a.nsub;
    const int nband = nsub == 140 ? 16 : 14;
    const bool sw = nsub == 112;
    int b = 0, end = mcica_band_ng(nsub, 0);
    while (ig >= end) end += mcica_band_ng(nsub, ++b);
    if (ig == 0) {
        for (int l = 0; l < nlay; ++l) {
            const size_t ix = (size_t)l * ncol + col;
            if (a.reicmcl) a.reicmcl[ix] = a.rei ? a.rei[ix] : 0.0;
            if (a.relqmcl) a.relqmcl[ix] = a.rel ? a.rel[ix] : 0.0;
        }
    }
    const size_t k0 = ((size_t)ig * ncol + col) * (icld == 3 ? 1 : nlay);
    double prev = 0.0, fprev = 0.0;
    for (int l = 0; l < nlay; ++l) {
        double x = (double)a.mt[k0 + (icld == 3 ? 0 : l)] / 4294967295.0;
        const size_t ix = (size_t)l * ncol + col, o = ix * nsub + ig, j = ix * nband + b;
        double f = a.cldfrac[ix];
        if (f < 1.e-20) f = 0.0;
        if (icld == 2 && l > 0) x = prev > 1.0 - fprev ? prev : x * (1.0 - fprev);
        prev = x;
        fprev = f;
        const bool cloudy = x >= 1.0 - f;
        __builtin_nontemporal_store(cloudy ? 1.0 : 0.0, a.cldfmcl + o);
        __builtin_nontemporal_store(cloudy ? a.tauc[j] : 0.0, a.taucmcl + o);
        if (a.ciwpmcl) __builtin_nontemporal_store(cloudy && a.ciwp ? a.ciwp[ix] : 0.0, a.ciwpmcl + o);
        if (a.clwpmcl) __builtin_nontemporal_store(cloudy && a.clwp ? a.clwp[ix] : 0.0, a.clwpmcl + o);
        if (sw) {
            __builtin_nontemporal_store(cloudy ? a.ssac[j] : 1.0, a.ssacmcl + o);
            __builtin_nontemporal_store(cloudy ? a.asmc[j] : 0.0, a.asmcmcl + o);
            __builtin_nontemporal_store(cloudy ? a.fsfc[j] : 0.0, a.fsfcmcl + o);
        }
    }
}

__global__ __launch_bounds__(256) void mcica_subcol_kernel(McicaSubcolArgs a, const McicaJump *__restrict__ jump,
                                                           const uint32_t *__restrict__ xs_jump)
{
#pragma clang fp contract(off)
    __shared__ uint32_t s_xs[KISS_XS_JUMPS * 32];
    for (int i = threadIdx.x; i < KISS_XS_JUMPS * 32; i += blockDim.x) s_xs[i] = xs_jump[i];
    __syncthreads();
    const int ncol = a.ncol, nlay = a.nlay, icld = a.icld, nsub = a.nsub;
    const int nband = nsub == 140 ? 16 : 14;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ncol * nsub) return;
    const int col = (int)(t / nsub), ig = (int)(t - (size_t)col * nsub);
    if (a.mt) {
        mcica_subcol_mt(a, col, ig);
        return;
    }
    if (a.play[col] * 1.e2 < a.play[(size_t)ncol + col] * 1.e2) {
        if (a.bad && ig == 0) atomicAdd(a.bad, 1);
        return;
    }
    // seeds (mcica_subcol_gen_lw.f90:330-333)
    uint32_t s[4];
    for (int k = 0; k < 4; ++k) {
        const double p = a.play[(size_t)k * ncol + col] * 1.e2;
        s[k] = (uint32_t)(int32_t)((p - (double)(int32_t)p) * 1000000000.0);
    }
    const McicaJump J = jump[ig];
    const uint64_t d = mcica_skip(a.permuteseed, icld, nlay, ig);
    s[0] = J.a1 * s[0] + J.c1;
    uint32_t v = s[1];
    for (int k = 0; k < KISS_XS_JUMPS && (d >> k); ++k) {
        if (!((d >> k) & 1u)) continue;
        uint32_t r = 0;
        for (int j = 0; j < 32; ++j)
            if ((v >> j) & 1u) r ^= s_xs[k * 32 + j];
        v = r;
    }
    s[1] = v;
    s[2] = mwc_jump(s[2], d, J.p3, MWC3_A, MWC3_M);
    s[3] = mwc_jump(s[3], d, J.p4, MWC4_A, MWC4_M);
    if (ig == 0) {
        for (int l = 0; l < nlay; ++l) {
            const size_t ix = (size_t)l * ncol + col;
            if (a.reicmcl) a.reicmcl[ix] = a.rei ? a.rei[ix] : 0.0;
            if (a.relqmcl) a.relqmcl[ix] = a.rel ? a.rel[ix] : 0.0;
        }
    }
    const int b = J.band;
    const bool sw = nsub == 112;
    const double r3 = icld == 3 ? kiss_step(s[0], s[1], s[2], s[3]) : 0.0;
    double prev = 0.0, fprev = 0.0;
    for (int l = 0; l < nlay; ++l) {
        double x = icld == 3 ? r3 : kiss_step(s[0], s[1], s[2], s[3]);
        const size_t ix = (size_t)l * ncol + col, o = ix * nsub + ig, j = ix * nband + b;
        double f = a.cldfrac[ix];
        if (f < 1.e-20) f = 0.0;
        if (icld == 2 && l > 0) x = prev > 1.0 - fprev ? prev : x * (1.0 - fprev);
        prev = x;
        fprev = f;
        const bool cloudy = x >= 1.0 - f;
        // streaming stores: the subcolumns are written once and read by a
        // later call
        __builtin_nontemporal_store(cloudy ? 1.0 : 0.0, a.cldfmcl + o);
        __builtin_nontemporal_store(cloudy ? a.tauc[j] : 0.0, a.taucmcl + o);
        if (a.ciwpmcl) __builtin_nontemporal_store(cloudy && a.ciwp ? a.ciwp[ix] : 0.0, a.ciwpmcl + o);
        if (a.clwpmcl) __builtin_nontemporal_store(cloudy && a.clwp ? a.clwp[ix] : 0.0, a.clwpmcl + o);
        if (sw) {
            __builtin_nontemporal_store(cloudy ? a.ssac[j] : 1.0, a.ssacmcl + o);
            __builtin_nontemporal_store(cloudy ? a.asmc[j] : 0.0, a.asmcmcl + o);
            __builtin_nontemporal_store(cloudy ? a.fsfc[j] : 0.0, a.fsfcmcl + o);
        }
    }
}

__global__ void mcica_zero_kernel(int *bad) { *bad = 0; }

hipError_t launch_mcica_subcol(const McicaSubcolArgs &a, const uint32_t *xs_jump, McicaJump *jump, hipStream_t s)
{
    if (a.bad) hipLaunchKernelGGL(mcica_zero_kernel, dim3(1), dim3(1), 0, s, a.bad);
    if (a.mt)
        hipLaunchKernelGGL(mt_stream_kernel, dim3(1), dim3(640), 0, s, (uint32_t)a.permuteseed,
                           mcica_mt_count(a.ncol, a.nlay, a.icld, a.nsub), a.mt);
    else
        hipLaunchKernelGGL(mcica_jump_kernel, dim3(1), dim3(192), 0, s, a.nlay, a.icld, a.permuteseed, a.nsub,
                           jump);
    const size_t n = (size_t)a.ncol * a.nsub;
    hipLaunchKernelGGL(mcica_subcol_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a,
                       (const McicaJump *)jump, xs_jump);
    return hipGetLastError();
}

}  // namespace rrtmg
