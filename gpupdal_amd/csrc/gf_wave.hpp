// gf_wave.hpp -- wavefront-wide primitives shared by the search kernels: a
// (squared distance, tie key) entry per lane, ascending bitonic sort across
// the 64 lanes, merge of two sorted lane-lists, inclusive scan.
#pragma once

#include "gf_internal.hpp"

namespace gf {

constexpr uint64_t kInfD = ~0ull;
constexpr uint32_t kInfU = ~0u;

struct Ent {
    uint64_t d;   // bit pattern of the non-negative FP64 squared distance
    uint32_t id;  // tie key between equal distances (point id or visit order)
    uint32_t j;   // payload (sorted position, or point id)
};

__device__ __forceinline__ Ent ent_inf() { return Ent{kInfD, kInfU, kInfU}; }

__device__ __forceinline__ bool ent_less(const Ent& a, const Ent& b)
{
    return a.d < b.d || (a.d == b.d && a.id < b.id);
}

__device__ __forceinline__ double ent_d2(const Ent& e)
{
    return __longlong_as_double(static_cast<long long>(e.d));
}

__device__ __forceinline__ Ent ent_shfl_xor(const Ent& e, int m)
{
    Ent r;
    r.d = __shfl_xor(static_cast<unsigned long long>(e.d), m);
    r.id = __shfl_xor(e.id, m);
    r.j = __shfl_xor(e.j, m);
    return r;
}

// Lane src's value, as a wave-uniform (scalar) value; src must be
// wave-uniform.  Unlike a shuffle, the compiler then knows every lane holds
// the same value, so branches on it stay scalar.
// (The builtins return int: each word goes through uint32_t, so that the low
// word is not sign-extended into the high one.)
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int src)
{
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), src));
}

__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int src)
{
    const uint32_t lo = lane_u32(static_cast<uint32_t>(v), src);
    const uint32_t hi = lane_u32(static_cast<uint32_t>(v >> 32), src);
    return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}

__device__ __forceinline__ Ent ent_lane(const Ent& e, int src)
{
    return Ent{lane_u64(e.d, src), lane_u32(e.id, src), lane_u32(e.j, src)};
}

// A value every lane already holds, as a wave-uniform (scalar) value.
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double uniform_f64(double v)
{
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(b))));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(b >> 32))));
    return __longlong_as_double(
        static_cast<long long>(static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32)));
}

__device__ __forceinline__ Ent ent_shfl(const Ent& e, int src)
{
    Ent r;
    r.d = __shfl(static_cast<unsigned long long>(e.d), src);
    r.id = __shfl(e.id, src);
    r.j = __shfl(e.j, src);
    return r;
}

// Ascending bitonic sort of one entry per lane across the wave.
__device__ __forceinline__ void bitonic_sort64(Ent& e, int lane)
{
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            Ent o = ent_shfl_xor(e, j);
            bool up = (lane & k) == 0;
            bool lower = (lane & j) == 0;
            bool take = (lower == up) ? ent_less(o, e) : ent_less(e, o);
            if (take) e = o;
        }
    }
}

// best (ascending) <- the 64 smallest of best U cand (cand ascending).
__device__ __forceinline__ void merge_sorted(Ent& best, const Ent& cand, int lane)
{
    Ent rev = ent_shfl(cand, 63 - lane);
    if (ent_less(rev, best)) best = rev;
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) {
        Ent o = ent_shfl_xor(best, j);
        bool lower = (lane & j) == 0;
        bool take = lower ? ent_less(o, best) : ent_less(best, o);
        if (take) best = o;
    }
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d);
        if (lane >= d) v += t;
    }
    return v;
}

} // namespace gf
