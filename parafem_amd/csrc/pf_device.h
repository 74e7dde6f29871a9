// Device helpers shared by the iteration kernels of libparafem_amd: state
// gating, wavefront and block reductions in the library's fixed order, the
// lane layout of the element products and the XCD-aware block order.
#pragma once

#include "pf_internal.h"

namespace {

constexpr int MAXG = PF_REDUCE_MAX_GRID;

enum { KIND_F64 = 0, KIND_F32 = 1, KIND_ACC = 2, KIND_WIDE = 3 };   // K_e storage / arithmetic
enum { RUN_ALWAYS = 0, RUN_LIVE = 1, RUN_REFINE = 2 };  // state gating

__device__ __forceinline__ bool skip(const pf_state *st, int gate)
{
    if (!st || gate == RUN_ALWAYS) return false;
    if (st->frozen) return true;
    return gate == RUN_REFINE && !st->refine;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = __dadd_rn(v, __shfl_down(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

// Folds NQ per-thread values of this block (sum or max per quantity) and
// stores the block results in partials[qid*MAXG + bidx].
template <int NQ>
__device__ void block_to_partials(double (&v)[NQ], const int (&qid)[NQ], const bool (&ismax)[NQ], double *partials,
                                  int bidx)
{
    __shared__ double ws[NQ][PF_BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = ismax[k] ? wave_max(v[k]) : wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NQ; ++k) ws[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            double s = ws[k][0];
            for (int w = 1; w < PF_BLOCK / 64; ++w) s = ismax[k] ? fmax(s, ws[k][w]) : __dadd_rn(s, ws[k][w]);
            partials[qid[k] * MAXG + bidx] = s;
        }
    }
}

// In a one-block kernel: folds the G block results of each quantity in the
// fixed order; out[] is valid in thread 0.
template <int NQ>
__device__ void fold_partials(double (&out)[NQ], const int (&qid)[NQ], const bool (&ismax)[NQ],
                              const double *partials, int G)
{
    __shared__ double ws[NQ][PF_BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        // Thread t adds entries t, t+256, ... in that order; the loads of a
        // batch of 16 are issued before the dependent additions.
        double s = 0.0;
        for (int b0 = threadIdx.x; b0 < G; b0 += 16 * PF_BLOCK) {
            double t[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int b = b0 + u * PF_BLOCK;
                t[u] = b < G ? partials[qid[k] * MAXG + b] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (b0 + u * PF_BLOCK < G) s = ismax[k] ? fmax(s, t[u]) : __dadd_rn(s, t[u]);
        }
        s = ismax[k] ? wave_max(s) : wave_sum(s);
        if (lane == 0) ws[k][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            double s = ws[k][0];
            for (int w = 1; w < PF_BLOCK / 64; ++w) s = ismax[k] ? fmax(s, ws[k][w]) : __dadd_rn(s, ws[k][w]);
            out[k] = s;
        }
    }
}

// Width of the lane group serving one element, elements per wavefront, and
// elements per lane group (EPL: a group loads EPL element matrices before it
// multiplies, to keep enough bytes in flight when an element is small).
template <int NT, int KIND> struct Lanes {
    static constexpr int W = NT > 32 ? 64 : (NT > 16 ? 32 : (NT > 8 ? 16 : 8));
    static constexpr int EPW = 64 / W;
    static constexpr int EPL = NT > 32 ? 1 : NT > 8 ? (KIND == 1 ? 2 : 1) : (KIND == 1 ? 8 : KIND == 0 ? 1 : 4);
    static constexpr int EPB = EPW * EPL * (PF_BLOCK / 64);   // elements per block
};

// Workgroups are dealt to the 8 XCDs round-robin by index.  Renumber them so
// that each XCD works on one contiguous range of logical blocks: neighbouring
// elements share nodes, and their gathers then hit the same L2.
__device__ __forceinline__ long xcd_block(int b, int nb)
{
    const int per = nb >> 3, rem = nb & 7, x = b & 7, i = b >> 3;
    return (long)x * per + (x < rem ? x : rem) + i;
}

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double bpermute64(int addr, double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)b);
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

}  // namespace
