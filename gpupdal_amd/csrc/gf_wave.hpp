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
