// Range-fused element products and scatter for the FP32-arithmetic (mixed)
// solver.
//
// The slot-list path writes every element product to utemp and sums each
// equation's slots in a second pass: utemp makes a round trip through HBM and
// the sums read 4-byte entries of other elements' rows.  Here a workgroup
// takes a range of R consecutive elements (4 hex20, 8 hex8 elastic, 32 hex8
// Laplace: one element per lane group), gathers p once per distinct equation
// of the range into LDS, forms the element products as the slot-list path
// does (packed symmetric FP32 hi, FP32 arithmetic on fl32(p)), sums them per
// equation in LDS (FP64) and adds each sum into u once.  Ranges that share an
// equation never run at the same time: the ranges are coloured (greedy, in
// range order) and each colour is one launch; the first colour to touch an
// equation writes u, the later ones add.  Inside a range the element rows are
// added in rounds of elements that share no equation, so every sum is formed
// in a fixed order: results are deterministic (bit-identical run to run and
// whatever else runs), though not in the slot-list path's order -- this path
// serves the mixed solver only, whose products are FP32 anyway.  p.u is
// summed per range over the range's equations (p(q) times its sum there) and
// folded over the ranges in range order; a fixed freedom (p123's penalty
// rows, p122's zero rows) is written by its first range as p*store (or 0),
// contributing p*u once.  On hex8 meshes the first range of each equation
// also forms the iteration's p = d + beta p there (the p update of the other
// paths), so p makes no separate pass and needs no FP32 copy.
//
// Reference: the product and SCATTER of p121.f90:89-97 (gather_scatter.f90:
// 694-850) and p123.f90:135-145 (fixed rows).
#include <algorithm>
#include <vector>

#include "pf_internal.h"
#include "pf_device.h"

namespace {

constexpr unsigned UE_FIRST = 1u << 31;   // this range is the first to touch the equation
constexpr unsigned UE_FIXED = 1u << 30;   // a fixed freedom (fixstore != 0)
constexpr unsigned UE_EQ = (1u << 30) - 1;
constexpr unsigned short LM_NONE = 0xFFFF;

template <int NT> struct RangeShape {
    static constexpr int R = Lanes<NT, KIND_F32>::EPW * (PF_BLOCK / 64);   // elements per range
    static_assert(R * NT <= PF_BLOCK, "one thread per distinct equation of a range");
};

// Whether the range kernels form p = d + beta p themselves (hex8: it saves
// the p pass and the FP32 copy of p that their gathers read).  For hex20 the
// separate pass is cheaper: forming p in the ranges (an extra d read and a
// scattered p write per equation) cost 0.11 ms per iteration at 1M elements,
// the pass 0.058 ms.
__host__ __device__ constexpr bool range_forms_p(int nt) { return nt <= 24; }

// One workgroup per range of this colour (ranges[blockIdx.x]).
template <int NT>
__global__ __launch_bounds__(PF_BLOCK) void k_ebe_range(
    const float *__restrict__ Khi, const unsigned short *__restrict__ lmap, const unsigned char *__restrict__ lround,
    int nround, const int *__restrict__ rptr, const unsigned *__restrict__ ue, const int *__restrict__ ranges,
    double *__restrict__ p, const double *__restrict__ d, double *__restrict__ u,
    const double *__restrict__ fixstore, int fix_zero, double *__restrict__ rpart, int nels,
    const pf_state *__restrict__ st)
{
    if (skip(st, RUN_LIVE)) return;
    const double beta = st->beta;
    const bool pnew = range_forms_p(NT) && !st->pfresh;
    using L = Lanes<NT, KIND_F32>;
    constexpr int R = RangeShape<NT>::R;
    constexpr int D = pf_sym_diags(NT), Q = pf_sym_quads(NT), RM = pf_sym_rem(NT), H = NT / 2;
    constexpr long PER = pf_sym_per(NT);
    __shared__ float pl[PF_BLOCK];                     // fl32(p) per distinct equation
    __shared__ double pacc[PF_BLOCK];                  // sums per distinct equation
    __shared__ __attribute__((aligned(16))) float ps[PF_BLOCK / 64][128];   // p_e twice per element
    const int r = ranges[blockIdx.x];
    const int u0 = rptr[r], nu = rptr[r + 1] - u0;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int sub = lane / L::W, li = lane % L::W;
    const int i = li < NT ? li : NT - 1;
    const int gb = sub * L::W;
    const long e = (long)r * R + wave * L::EPW + sub;
    const bool ve = e < nels;
    const long ec = ve ? e : (long)nels - 1;
    // Every load that depends on nothing else first -- the range's equation
    // list, the lane's local index and round, the element matrix -- then
    // the ones that need the equation numbers (p, and u where this range
    // adds to it): two memory latencies before the products.
    const unsigned uf = t < nu ? ue[u0 + t] : 0u;
    const unsigned short k = (ve && li < NT) ? lmap[ec * NT + i] : LM_NONE;
    const int rd = ve ? lround[e] : 255;
    float kd[D];
    {
        const float *Ke = Khi + ec * PER;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const f4_t v = __builtin_nontemporal_load(reinterpret_cast<const f4_t *>(Ke) + q * NT + i);
            kd[4 * q] = v.x; kd[4 * q + 1] = v.y; kd[4 * q + 2] = v.z; kd[4 * q + 3] = v.w;
        }
#pragma unroll
        for (int m = 0; m < RM; ++m) kd[4 * Q + m] = __builtin_nontemporal_load(Ke + 4 * Q * NT + RM * i + m);
        kd[H] = __builtin_nontemporal_load(Ke + H * NT + i % H);
    }
    const int eq = (int)(uf & UE_EQ);
    double pd = 0.0, uold = 0.0, fsv = 0.0;
    if (t < nu) {
        // every load before any result is used (one memory latency); then
        // p = d + beta p (the reference's p update, p121.f90:102) by the
        // equation's first range, stored with u at the end (a store here
        // would hold the barrier below until it completes); the later
        // colours read it
        if (!(uf & UE_FIXED) && !(uf & UE_FIRST)) uold = u[eq];
        if ((uf & UE_FIXED) && (uf & UE_FIRST) && !fix_zero) fsv = fixstore[eq];
        const double pold = p[eq];
        const bool form = pnew && (uf & UE_FIRST);
        const double dq = form ? d[eq] : 0.0;
        pd = form ? __dadd_rn(dq, __dmul_rn(pold, beta)) : pold;
        pl[t] = (float)pd;
        pacc[t] = 0.0;
    }
    __syncthreads();
    // the element product of this lane group's element (as the slot-list
    // path's FP32 kind forms it)
    const float pv = k != LM_NONE ? pl[k] : 0.0f;
    float *pg = &ps[wave][0] + 2 * gb;
    if (li < NT) {
        pg[li] = pv;
        pg[li + NT] = pv;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) acc = __fmaf_rn(kd[d], pg[i + d], acc);
    int j = i;   // (i - d) mod NT: the lane sending K(i-d, i) p(i)
#pragma unroll
    for (int d = 1; d < D - 1; ++d) {
        j = j == 0 ? NT - 1 : j - 1;
        const float tt = __fmul_rn(kd[d], pv);
        acc = __fadd_rn(acc, __int_as_float(__builtin_amdgcn_ds_bpermute((gb + j) << 2, __float_as_int(tt))));
    }
    // per-equation sums: rounds of elements that share no equation
    for (int q = 0; q < nround; ++q) {
        if (rd == q && k != LM_NONE) pacc[k] = __dadd_rn(pacc[k], (double)acc);
        __syncthreads();
    }
    double dot[1] = {0.0};
    if (t < nu) {
        if (pnew && (uf & UE_FIRST)) p[eq] = pd;
        const double s = pacc[t];
        if (uf & UE_FIXED) {
            if (uf & UE_FIRST) {
                const double v = fix_zero ? 0.0 : __dmul_rn(pd, fsv);
                u[eq] = v;
                dot[0] = __dmul_rn(pd, v);
            }
        } else {
            u[eq] = (uf & UE_FIRST) ? s : __dadd_rn(uold, s);
            dot[0] = __dmul_rn(pd, s);
        }
    }
    constexpr int qid[1] = {0};
    constexpr bool mx[1] = {false};
    block_to_partials<1>(dot, qid, mx, rpart, r);
}

// p.u from the per-range partials in two fixed-order stages: block b of
// NF folds partials b, b + NF, ... (stride order, then the block tree) into
// fpart[b]; one block folds the NF of them and forms alpha.
constexpr int NF = 256;
__global__ __launch_bounds__(PF_BLOCK) void k_fold_ranges1(const double *__restrict__ rpart, int nrange,
                                                          double *__restrict__ fpart, const pf_state *st)
{
    if (skip(st, RUN_LIVE)) return;
    double v[1] = {0.0};
    for (int b = blockIdx.x + NF * threadIdx.x; b < nrange; b += NF * PF_BLOCK) v[0] = __dadd_rn(v[0], rpart[b]);
    constexpr int qid[1] = {0};
    constexpr bool mx[1] = {false};
    block_to_partials<1>(v, qid, mx, fpart, blockIdx.x);
}

__global__ __launch_bounds__(PF_BLOCK) void k_fold_alpha_ranges(const double *fpart, pf_state *st)
{
    if (skip(st, RUN_LIVE)) return;
    constexpr int qid[1] = {0};
    constexpr bool mx[1] = {false};
    double out[1];
    fold_partials<1>(out, qid, mx, fpart, NF);
    if (threadIdx.x == 0) {
        const double alpha = st->up / out[0];
        st->alpha = alpha;
        if (!isfinite(alpha)) {
            st->breakdown = 1;
            st->frozen = 1;
        }
    }
}

template <int NT>
void launch_ranges_nt(pf_ctx *c)
{
    const pf_ranges &F = *c->rng;
    for (int col = 0; col < F.ncol; ++col) {
        const int nb = F.cstart[col + 1] - F.cstart[col];
        if (nb <= 0) continue;
        hipLaunchKernelGGL(k_ebe_range<NT>, dim3(nb), dim3(PF_BLOCK), 0, c->stream, c->K32hi, F.lmap, F.lround,
                           F.nround, F.rptr, F.ue, F.cperm + F.cstart[col], c->p, c->d, c->u, c->fixstore,
                           c->fix_zero ? 1 : 0, F.rpart, c->nels, c->state);
    }
}

// ---- building the ranges on the device ------------------------------------
//
// A workgroup per range sorts the range's slots as keys equation << 8 | slot
// (slot = position among the range's R*NT slots): equal equations become
// neighbours, in slot order.  k_rng_count counts the distinct equations;
// after a scan of the counts k_rng_fill writes each range's ascending list,
// the local index of every slot (the rank of its equation in the list) and
// the rounds.  The colours of the ranges are chosen on the host, greedily in
// range order (each sequentially depends on the ranges before it), from the
// lists copied back; k_rng_mask and k_rng_flags then mark the first range
// (lowest colour) of every equation and the fixed freedoms.
static_assert(PF_BLOCK == 256, "a range's slot index takes the low 8 bits of a key");
constexpr unsigned long long KEY_NONE = ~0ull;

template <int NT>
__device__ __forceinline__ void range_keys_sorted(const int *__restrict__ gg, long nslot, long r,
                                                  unsigned long long *key)
{
    constexpr int W = RangeShape<NT>::R * NT;
    const int t = threadIdx.x;
    const long s = r * W + t;
    unsigned long long kv = KEY_NONE;
    if (t < W && s < nslot) {
        const int g = gg[s];
        if (g > 0) kv = ((unsigned long long)(unsigned)g << 8) | (unsigned)t;
    }
    key[t] = kv;
    __syncthreads();
    // bitonic sort of PF_BLOCK keys, ascending
    for (int k = 2; k <= PF_BLOCK; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            const int x = t ^ j;
            if (x > t) {
                const unsigned long long a = key[t], b = key[x];
                if ((a > b) == ((t & k) == 0)) {
                    key[t] = b;
                    key[x] = a;
                }
            }
            __syncthreads();
        }
}

template <int NT>
__global__ __launch_bounds__(PF_BLOCK) void k_rng_count(const int *__restrict__ gg, long nslot,
                                                       int *__restrict__ cnt, int *__restrict__ err)
{
    __shared__ unsigned long long key[PF_BLOCK];
    const long r = blockIdx.x;
    range_keys_sorted<NT>(gg, nslot, r, key);
    const int t = threadIdx.x;
    const unsigned long long kv = key[t], prev = t ? key[t - 1] : KEY_NONE;
    const bool valid = kv != KEY_NONE;
    const bool repeat = valid && t && (prev >> 8) == (kv >> 8);
    // the same equation in two slots of one element
    if (repeat && (int)(prev & 0xFF) / NT == (int)(kv & 0xFF) / NT) atomicOr(err, 1);
    const int n = __syncthreads_count(valid && !repeat);
    if (t == 0) cnt[r] = n;
}

template <int NT>
__global__ __launch_bounds__(PF_BLOCK) void k_rng_fill(const int *__restrict__ gg, long nslot, long nels,
                                                      const int *__restrict__ rptr, unsigned *__restrict__ ue,
                                                      unsigned short *__restrict__ lmap,
                                                      unsigned char *__restrict__ lround, int *__restrict__ nround)
{
    constexpr int R = RangeShape<NT>::R, W = R * NT;
    static_assert(R <= 32, "rounds are bits of a 32-bit mask");
    __shared__ unsigned long long key[PF_BLOCK];
    __shared__ unsigned short lm[PF_BLOCK];
    __shared__ unsigned rmask[PF_BLOCK];
    __shared__ int wsum[PF_BLOCK / 64];
    __shared__ unsigned busy;
    const long r = blockIdx.x;
    range_keys_sorted<NT>(gg, nslot, r, key);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const unsigned long long kv = key[t];
    const bool valid = kv != KEY_NONE;
    const bool first = valid && (t == 0 || (key[t - 1] >> 8) != (kv >> 8));
    const unsigned long long bal = __ballot(first);
    const int upto = __popcll(bal & ((2ull << lane) - 1ull));   // lanes 0..lane (all 64 at lane 63)
    if (lane == 63) wsum[wave] = __popcll(bal);
    lm[t] = LM_NONE;
    rmask[t] = 0u;
    __syncthreads();
    int grp = upto - 1;   // rank of this key's equation among the range's distinct ones
    for (int w = 0; w < wave; ++w) grp += wsum[w];
    if (first) ue[rptr[r] + grp] = (unsigned)(kv >> 8);
    if (valid) lm[kv & 0xFF] = (unsigned short)grp;
    __syncthreads();
    if (t < W && r * W + t < nslot) lmap[r * W + t] = lm[t];
    // rounds: each element takes the lowest round that no earlier element of
    // the range sharing an equation with it has taken (at most e bits are
    // set before element e, so a round below 32 is free)
    const long e0 = r * R;
    int top = 0;
    for (int e = 0; e < R && e0 + e < nels; ++e) {
        if (t == 0) busy = 0u;
        __syncthreads();
        const unsigned short k = t < NT ? lm[e * NT + t] : LM_NONE;
        if (k != LM_NONE) atomicOr(&busy, rmask[k]);
        __syncthreads();
        const int rd = __builtin_ctz(~busy);
        if (k != LM_NONE) rmask[k] |= 1u << rd;
        if (t == 0) {
            lround[e0 + e] = (unsigned char)rd;
            top = max(top, rd + 1);
        }
        __syncthreads();
    }
    if (t == 0) atomicMax(nround, top);
}

// mask[q] |= 1 << colour of every range listing q.
__global__ __launch_bounds__(PF_BLOCK) void k_rng_mask(const int *__restrict__ rptr, const unsigned *__restrict__ ue,
                                                      const int *__restrict__ rcol, unsigned *__restrict__ mask)
{
    const long r = blockIdx.x;
    const unsigned bit = 1u << rcol[r];
    for (int a = rptr[r] + threadIdx.x; a < rptr[r + 1]; a += PF_BLOCK) atomicOr(&mask[ue[a]], bit);
}

__global__ __launch_bounds__(PF_BLOCK) void k_rng_flags(const int *__restrict__ rptr, unsigned *__restrict__ ue,
                                                       const int *__restrict__ rcol, const unsigned *__restrict__ mask,
                                                       const double *__restrict__ fixstore)
{
    const long r = blockIdx.x;
    const int col = rcol[r];
    for (int a = rptr[r] + threadIdx.x; a < rptr[r + 1]; a += PF_BLOCK) {
        const unsigned q = ue[a];
        unsigned v = q;
        if (__builtin_ctz(mask[q]) == col) v |= UE_FIRST;
        if (fixstore && fixstore[q] != 0.0) v |= UE_FIXED;
        ue[a] = v;
    }
}

template <int NT>
void launch_rng_count(pf_ctx *c, long nrange, long nslot, int *cnt, int *err)
{
    hipLaunchKernelGGL(k_rng_count<NT>, dim3(nrange), dim3(PF_BLOCK), 0, c->stream, c->g_g, nslot, cnt, err);
}

template <int NT>
void launch_rng_fill(pf_ctx *c, long nrange, long nslot, const pf_ranges &F, int *nround)
{
    hipLaunchKernelGGL(k_rng_fill<NT>, dim3(nrange), dim3(PF_BLOCK), 0, c->stream, c->g_g, nslot, (long)c->nels,
                       F.rptr, F.ue, F.lmap, F.lround, nround);
}

template <class T>
void free_dev(pf_ctx *c, T *&p, size_t n)
{
    pf_free(c, p, n * sizeof(T));
    p = nullptr;
}

}  // namespace

void pf_ranges_drop(pf_ctx *c)
{
    if (!c->rng) return;
    pf_ranges &F = *c->rng;
    (void)hipStreamSynchronize(c->stream);
    free_dev(c, F.lmap, (size_t)F.nslot);
    free_dev(c, F.lround, (size_t)F.nels);
    free_dev(c, F.rptr, (size_t)F.nrange + 1);
    free_dev(c, F.ue, (size_t)F.nue);
    free_dev(c, F.cperm, (size_t)F.nrange);
    free_dev(c, F.rpart, (size_t)F.nrange + NF);
    delete c->rng;
    c->rng = nullptr;
}

// Builds the ranges of the current steering and fixed rows: distinct
// equations per range, the local index of every element slot and the rounds
// inside a range on the device; the colours of the ranges on the host from
// the lists (greedy in range order); the first range of every equation and
// the fixed freedoms on the device again.  Returns PF_ERR_ARG when the
// steering does not fit the scheme (an element listing an equation twice, or
// more than 32 colours): the solve then keeps the slot-list scatter.
int pf_ranges_build(pf_ctx *c)
{
    pf_ranges_drop(c);
    const int NT = c->ntot;
    int R = 0;
    switch (NT) {
    case 60: R = RangeShape<60>::R; break;
    case 24: R = RangeShape<24>::R; break;
    case 8: R = RangeShape<8>::R; break;
    default: return PF_ERR_ARG;
    }
    const long nels = c->nels, nslot = (long)NT * nels;
    const long nrange = (nels + R - 1) / R;
    if (nrange >= INT32_MAX) return PF_ERR_ARG;
    c->rng = new pf_ranges();
    pf_ranges &F = *c->rng;
    F.nslot = nslot;
    F.nels = nels;
    F.nrange = nrange;
    const int nb = pf_scan_blocks(nrange);
    int *cnt = nullptr, *bsum = nullptr, *flags = nullptr, *rcol = nullptr;
    unsigned *mask = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(c->stream);
        free_dev(c, cnt, (size_t)nrange);
        free_dev(c, bsum, (size_t)nb);
        free_dev(c, flags, 4);
        free_dev(c, rcol, (size_t)nrange);
        free_dev(c, mask, (size_t)c->neq + 1);
    };
    auto fail = [&](int rc) {
        release();
        pf_ranges_drop(c);
        return rc;
    };
    int rc = pf_alloc(c, (void **)&cnt, (size_t)nrange * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&bsum, (size_t)nb * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&flags, 4 * sizeof(int));   // error, list entries, rounds, zero counts
    if (!rc) rc = pf_alloc(c, (void **)&F.rptr, ((size_t)nrange + 1) * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&F.lmap, (size_t)nslot * sizeof(unsigned short));
    if (!rc) rc = pf_alloc(c, (void **)&F.lround, (size_t)nels);
    if (rc) return fail(rc);
    int host[3] = {0, 0, 0};
    hipError_t he = hipMemsetAsync(flags, 0, 4 * sizeof(int), c->stream);
    if (he == hipSuccess) he = hipMemsetAsync(F.rptr, 0, sizeof(int), c->stream);
    if (he == hipSuccess) {
        switch (NT) {
        case 60: launch_rng_count<60>(c, nrange, nslot, cnt, flags); break;
        case 24: launch_rng_count<24>(c, nrange, nslot, cnt, flags); break;
        default: launch_rng_count<8>(c, nrange, nslot, cnt, flags); break;
        }
        pf_scan_launch(c, cnt, (int)nrange, bsum, F.rptr, flags + 1, flags + 3);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipMemcpyAsync(host, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) return fail(pf_set_hip_error(he, "range lists", __FILE__, __LINE__));
    if (host[0]) return fail(PF_ERR_ARG);   // an equation twice in one element
    F.nue = host[1];
    rc = pf_alloc(c, (void **)&F.ue, (size_t)std::max<long>(F.nue, 1) * sizeof(unsigned));
    if (!rc) rc = pf_alloc(c, (void **)&rcol, (size_t)nrange * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&mask, ((size_t)c->neq + 1) * sizeof(unsigned));
    if (!rc) rc = pf_alloc(c, (void **)&F.cperm, (size_t)nrange * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&F.rpart, ((size_t)nrange + NF) * sizeof(double));
    if (rc) return fail(rc);
    switch (NT) {
    case 60: launch_rng_fill<60>(c, nrange, nslot, F, flags + 2); break;
    case 24: launch_rng_fill<24>(c, nrange, nslot, F, flags + 2); break;
    default: launch_rng_fill<8>(c, nrange, nslot, F, flags + 2); break;
    }
    he = hipGetLastError();
    // the lists for the colouring
    std::vector<int> rptr(nrange + 1);
    std::vector<unsigned> ue((size_t)F.nue);
    if (he == hipSuccess)
        he = hipMemcpyAsync(rptr.data(), F.rptr, rptr.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess && F.nue)
        he = hipMemcpyAsync(ue.data(), F.ue, ue.size() * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(host + 2, flags + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) return fail(pf_set_hip_error(he, "range lists", __FILE__, __LINE__));
    F.nround = std::max(host[2], 1);
    // colours: a range takes the lowest colour no earlier range sharing an
    // equation has taken
    std::vector<int> col(nrange);
    std::vector<unsigned> used_by((size_t)c->neq + 1, 0u);
    int ncol = 1;
    for (long r = 0; r < nrange; ++r) {
        unsigned used = 0;
        for (int a = rptr[r]; a < rptr[r + 1]; ++a) used |= used_by[ue[a]];
        if (~used == 0u) return fail(PF_ERR_ARG);   // more than 32 colours
        const int k = __builtin_ctz(~used);
        col[r] = k;
        ncol = std::max(ncol, k + 1);
        for (int a = rptr[r]; a < rptr[r + 1]; ++a) used_by[ue[a]] |= 1u << k;
    }
    // ranges by colour, ascending range order inside a colour
    std::vector<int> cstart(ncol + 1, 0), cperm(nrange);
    for (long r = 0; r < nrange; ++r) cstart[col[r] + 1]++;
    for (int k = 0; k < ncol; ++k) cstart[k + 1] += cstart[k];
    {
        std::vector<int> fill(cstart.begin(), cstart.end() - 1);
        for (long r = 0; r < nrange; ++r) cperm[fill[col[r]]++] = (int)r;
    }
    F.ncol = ncol;
    F.cstart = cstart;
    he = hipMemcpyAsync(rcol, col.data(), nrange * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(F.cperm, cperm.data(), nrange * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess) he = hipMemsetAsync(mask, 0, ((size_t)c->neq + 1) * sizeof(unsigned), c->stream);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(k_rng_mask, dim3(nrange), dim3(PF_BLOCK), 0, c->stream, F.rptr, F.ue, rcol, mask);
        hipLaunchKernelGGL(k_rng_flags, dim3(nrange), dim3(PF_BLOCK), 0, c->stream, F.rptr, F.ue, rcol, mask,
                           c->fixstore);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) return fail(pf_set_hip_error(he, "range flags", __FILE__, __LINE__));
    release();
    return PF_OK;
}

// The mixed iteration's product, sums into u and alpha (replaces k_ebe,
// k_scatter_pu and k_fold_alpha).
int pf_ranges_iteration(pf_ctx *c)
{
    if (!c->rng) return PF_ERR_STATE;
    switch (c->ntot) {
    case 60: launch_ranges_nt<60>(c); break;
    case 24: launch_ranges_nt<24>(c); break;
    default: launch_ranges_nt<8>(c); break;
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}

bool pf_ranges_form_p(const pf_ctx *c) { return range_forms_p(c->ntot); }

int pf_ranges_fold(pf_ctx *c)
{
    double *fpart = c->rng->rpart + c->rng->nrange;   // NF more doubles past the range partials
    hipLaunchKernelGGL(k_fold_ranges1, dim3(NF), dim3(PF_BLOCK), 0, c->stream, c->rng->rpart,
                       (int)c->rng->nrange, fpart, c->state);
    hipLaunchKernelGGL(k_fold_alpha_ranges, dim3(1), dim3(PF_BLOCK), 0, c->stream, fpart, c->state);
    PF_HIP(hipGetLastError());
    return PF_OK;
}
