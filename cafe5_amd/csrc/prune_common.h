// prune_common.h -- device helpers shared by the pruning kernels: read-only
// tables through the scalar cache, the FP32 flush-bound constants and the
// root reduction (base_model.cpp:80-90 takes max_j, gamma_core.cpp:154-161
// the same maximum of x_j prior_j).
#pragma once
#include "engine.h"

namespace cafe {

// Read-only tables of a launch (program steps, tiles, buckets, node -> matrix
// slots) read through the constant address space: uniform loads of them go
// through the scalar cache and do not count against the vector-memory
// counter that the prefetched operands are waited on.  The kernels never
// write them; earlier kernels that do (the recompute-list packing) have
// finished before a launch starts.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_const(const T* q)
{
    return (const __attribute__((address_space(4))) T*)q;
}

// Element i of a read-only table of 16-byte-multiple records, via as_const.
template <typename T>
__device__ __forceinline__ T ld_const(const T* q, long long i)
{
    typedef int int4v __attribute__((ext_vector_type(4)));
    static_assert(sizeof(T) % sizeof(int4v) == 0, "records are whole 16-byte units");
    constexpr int n = sizeof(T) / sizeof(int4v);
    const __attribute__((address_space(4))) int4v* c = as_const((const int4v*)(q + i));
    struct Words { int4v w[n]; } r;
#pragma unroll
    for (int k = 0; k < n; ++k) r.w[k] = c[k];
    return __builtin_bit_cast(T, r);
}

constexpr double kLn2 = 0.69314718055994530942;
// FP32 flush threshold (smallest normal float): the flush bound charges this
// for every quantity FP32 could have flushed to zero.
constexpr double kTau32 = 1.1754943508222875e-38;
constexpr double kTwoPowMinusMatExp = 1.0 / (double)(1 << kMatExp);
// an FP32 likelihood is accepted when the flush bound is below this fraction
constexpr double kCertEta = 1e-9;
// below this, FP64 values may lose bits to subnormals; such families are
// recomputed in the reference's summation order (PLAIN64)
constexpr double kFp64Floor = 1e-290;

// 1/mant rounded up (mant in [0.5, 1)): the FP32 reciprocal is within one
// ulp, and the factor covers that and the rounding of mant to FP32.
__device__ __forceinline__ double rcp_up(float mant)
{
    return (double)__builtin_amdgcn_rcpf(mant) * (1.0 + 0x1p-20);
}

// Running root reduction of one lane: best key over its rows.
struct RootBest {
    double key;   // x * prior (scaled path: in units of 2^e_sum)
    double x;     // x at the best key
    int j;        // its index (smallest j on ties)
    double xmax;  // max_j x_j
};

__device__ __forceinline__ void root_take(RootBest& b, double key, double x, int j)
{
    if (b.j < 0 || b.key < key || (b.key == key && j < b.j)) {
        b.key = key;
        b.x = x;
        b.j = j;
    }
    b.xmax = b.xmax > x ? b.xmax : x;
}

__device__ __forceinline__ void root_merge(RootBest& b, double key, double x, int j, double xmax)
{
    if (j >= 0) {
        const double xm = b.xmax > xmax ? b.xmax : xmax;
        root_take(b, key, x, j);
        b.xmax = xm;
    }
}

}  // namespace cafe
