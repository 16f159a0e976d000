// flx_common.h -- shared definitions for the flx HIP kernels (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cfloat>
#include <type_traits>

#define FLX_DEV static __device__ __forceinline__

// Compile-time loop: f(std::integral_constant<int, I>) for I = B .. E-1.
// Bodies that index register arrays by the loop variable need it constant;
// a loop the compiler leaves rolled indexes registers through
// s_set_gpr_idx, one serialised move per access.
template <int B, int E>
struct StaticFor {
    template <typename F>
    __device__ __forceinline__ static void run(F &&f)
    {
        if constexpr (B < E) {
            f(std::integral_constant<int, B>{});
            StaticFor<B + 1, E>::run(f);
        }
    }
};

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Status codes returned by every flx_* entry point.
enum flx_status_internal {
    FLX_OK = 0,
    FLX_ERR_ARG = 1,        // invalid argument (negative size, bad ld, ...)
    FLX_ERR_HIP = 2,        // a HIP runtime call failed
    FLX_ERR_ALLOC = 3,      // device allocation failed
    FLX_ERR_SYNC = 4,       // a bounded device-side wait gave up
    FLX_ERR_UNSUPPORTED = 5 // shape or option outside what the kernels cover
};

// Agent-scope relaxed atomics: coherent across the XCDs of one device
// without fences (each 64-bit word is single-copy atomic).
// Latency-bound kernels on the factorisation's critical path raise their
// waves' instruction-issue priority (s_setprio, per wave, within the compute
// unit), so a concurrent trailing update sharing the compute units does not
// stretch their dependency chains.
FLX_DEV void critical_path_priority() { __builtin_amdgcn_s_setprio(3); }

FLX_DEV uint64_t ld_agent(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
FLX_DEV void st_agent(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A double posted as two 64-bit words, each tagged with a 32-bit sequence
// number in its high half, so a reader can verify every word it reads
// belongs to the step it expects without any fence or flag ordering.
FLX_DEV void post_double(uint64_t *w, double v, uint32_t seq)
{
    uint64_t b = __double_as_longlong(v);
    st_agent(w, ((uint64_t)seq << 32) | (b >> 32));
    st_agent(w + 1, ((uint64_t)seq << 32) | (b & 0xffffffffull));
}
FLX_DEV double join_double(uint64_t hi, uint64_t lo)
{
    return __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) |
                                            (lo & 0xffffffffull)));
}
FLX_DEV uint32_t seq_of(uint64_t w) { return (uint32_t)(w >> 32); }

// Wave-level (64 lanes) reductions through cross-lane shuffles.
FLX_DEV double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
FLX_DEV double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Pivot candidate ordering: larger |value| wins, ties go to the lower row
// (cxxblas/level1/iamax.tcc:42-58 keeps the first maximum).
FLX_DEV bool cand_better(double a, int ia, double b, int ib)
{
    return (a > b) || (a == b && ia < ib);
}

// Full-wave unsigned max / min through DPP (row_shr 1,2,3,4,8 then
// row_bcast 15/31); the result is read from lane 63 and is uniform.
FLX_DEV uint32_t wave_umax(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x113, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xe, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xc, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
FLX_DEV uint32_t wave_umin(uint32_t v)
{
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x111, 0xf, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x112, 0xf, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x113, 0xf, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x114, 0xf, 0xe, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x118, 0xf, 0xc, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x142, 0xa, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x143, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// max of non-negative doubles through their bit patterns
FLX_DEV void atomic_max_nonneg(unsigned long long *slot, double v)
{
    if (v > 0.0 || v != v) atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

// power of two 2^-e with v = f 2^e, f in [0.5, 1): scales v into [0.5, 1)
FLX_DEV double pow2_recip(double v)
{
    if (!(v > 0.0) || !(v <= DBL_MAX)) return 1.0;
    int e;
    frexp(v, &e);
    return ldexp(1.0, -e);
}
