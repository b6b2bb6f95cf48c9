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

}  // namespace prism
