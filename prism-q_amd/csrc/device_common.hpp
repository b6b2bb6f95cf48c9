// device_common.hpp -- device helpers shared by the MD kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace prism {

// The wave's index in its workgroup, as a value the compiler knows to be the
// same on every lane: from threadIdx.x >> 6 it cannot tell, so whatever is
// derived from it (a system index, its descriptor and state, a unit code)
// would be read with vector loads and branched on under the exec mask rather
// than with scalar loads and scalar branches.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ inline int wave_min_int(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Orders this wavefront's LDS accesses (cross-lane LDS communication inside
// one wave without a workgroup barrier).
__device__ inline void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// float -> int64 fixed point, saturated so a pathological contribution cannot
// wrap the accumulator.
__device__ inline long long to_fixed(float v, int bits) {
    double d = (double)v * (double)(1ll << bits);
    d = fmin(fmax(d, -4.0e18), 4.0e18);
    return (long long)rint(d);
}
__device__ inline void add_fixed(long long *acc, float v, int bits) {
    atomicAdd((unsigned long long *)acc, (unsigned long long)to_fixed(v, bits));
}
__device__ inline float from_fixed(long long v, int bits) {
    return (float)((double)v * (1.0 / (double)(1ll << bits)));
}

// Philox4x32-10 (Salmon et al., SC'11), Random123 constants.
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// Three standard normals of the Langevin stream: key (seed low word, stream),
// counter (atom, step low word, step bits 32..47 << 16 | purpose << 8, 0);
// Box-Muller with u1 = (r0 + 1) 2^-32, u2 = r1 2^-32 (and r2, r3).
__device__ inline void langevin_normals(uint32_t seed_lo, uint32_t stream, uint32_t atom, unsigned long long step,
                                        uint32_t purpose, float out[3]) {
    uint32_t c[4] = {atom, (uint32_t)step, (uint32_t)(((step >> 32) & 0xFFFFull) << 16) | ((purpose & 0xFFu) << 8), 0u};
    philox4x32_10(c, seed_lo, stream);
    // Box-Muller on the hardware transcendentals: v_log_f32 is log2, and
    // v_sin_f32 / v_cos_f32 take the angle in revolutions, which is exactly
    // the uniform u2 of the transform (angle 2 pi u2).
    const float s = 2.3283064365386963e-10f;  // 2^-32
    const float a = sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(fmaf((float)c[0], s, s)));  // -2 ln u1
    const float b = sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(fmaf((float)c[2], s, s)));
    const float t1 = (float)c[1] * s, t3 = (float)c[3] * s;
    out[0] = a * __builtin_amdgcn_cosf(t1);
    out[1] = a * __builtin_amdgcn_sinf(t1);
    out[2] = b * __builtin_amdgcn_cosf(t3);
}

// Director schedule (protocol_director.cu:58-123): temperature table and the
// equilibration friction ramp at schedule step `step`.
__device__ inline void schedule(const SysDesc &D, unsigned long long step, float &T, float &gamma) {
    // every field is read before the first comparison, so the loads issue
    // together instead of one per branch taken
    const long long s = (long long)step;
    const long long e_cold = D.cold_hold_end, e_ramp = D.ramp_end, e_warm = D.warm_hold_end, e_down = D.ramp_down_end;
    const long long e_eq = D.equilibration_steps;
    const float t0 = D.t_start, t1 = D.t_end, g_base = D.gamma_base, g_eq = D.gamma_eq;
    const bool warm_tail = D.warm_tail;
    if (s < e_cold) T = t0;
    else if (s < e_ramp) T = t0 + (float)(s - e_cold) / (float)(e_ramp - e_cold) * (t1 - t0);
    else if (s < e_warm) T = t1;
    else if (s < e_down) T = t1 - (float)(s - e_warm) / (float)(e_down - e_warm) * (t1 - t0);
    else T = warm_tail ? t1 : t0;
    gamma = g_base;
    if (s < e_eq) {
        const float decay = expf(-3.0f * (float)s / (float)e_eq);
        gamma = g_eq * decay + g_base * (1.0f - decay);
    }
}

// Value of lane (l - R) mod 16 of the same 16-lane row: DPP row_ror, no LDS
// traffic.  Four rotations by 1, 2, 4, 8 leave every lane of a row with the
// row's minimum or maximum.
template <int R> __device__ inline float row_ror(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + R, 0xf, 0xf, false));
}
__device__ inline float row_min(float v) {
    v = fminf(v, row_ror<1>(v));
    v = fminf(v, row_ror<2>(v));
    v = fminf(v, row_ror<4>(v));
    return fminf(v, row_ror<8>(v));
}
__device__ inline float row_max(float v) {
    v = fmaxf(v, row_ror<1>(v));
    v = fmaxf(v, row_ror<2>(v));
    v = fmaxf(v, row_ror<4>(v));
    return fmaxf(v, row_ror<8>(v));
}

// The current bounding box of each quarter (16 sorted slots) of block b, the
// box the nonbonded kernel culls j-atoms against, and of the whole block, the
// box its tile skip tests: one wave, lane l on slot b*64 + l, so the four
// 16-lane rows are the four quarters.  Padding slots take no part; a quarter
// of padding only gets an empty box (half extent -3e38), which no j-atom
// reaches.  Stored as (centre, half extent) pairs, the block's after its
// quarters' (QBOX_PER_BLOCK pairs per block).
__device__ inline void quarter_boxes(const DevPtrs &P, int b, int lane) {
    const int slot = b * BLOCK + lane;
    const float4 x = P.xq[slot];
    const bool real = P.perm[slot] >= 0;
    float lo[3] = {real ? x.x : 3.0e38f, real ? x.y : 3.0e38f, real ? x.z : 3.0e38f};
    float hi[3] = {real ? x.x : -3.0e38f, real ? x.y : -3.0e38f, real ? x.z : -3.0e38f};
    for (int k = 0; k < 3; ++k) {
        lo[k] = row_min(lo[k]);
        hi[k] = row_max(hi[k]);
    }
    float4 *q = P.qbox + (size_t)b * QBOX_PER_BLOCK * 2;
    if ((lane & (SUB - 1)) == 0) {
        q[lane / SUB * 2] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), 0.0f);
        q[lane / SUB * 2 + 1] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]), 0.5f * (hi[2] - lo[2]), 0.0f);
    }
    for (int k = 0; k < 3; ++k) {  // across the four rows
        lo[k] = fminf(lo[k], __shfl_xor(lo[k], 16));
        lo[k] = fminf(lo[k], __shfl_xor(lo[k], 32));
        hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], 16));
        hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], 32));
    }
    if (lane == 0) {
        q[QUARTERS * 2] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), 0.0f);
        q[QUARTERS * 2 + 1] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]), 0.5f * (hi[2] - lo[2]), 0.0f);
    }
}

}  // namespace prism
