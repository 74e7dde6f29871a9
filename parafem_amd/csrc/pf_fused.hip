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
// contributing p*u once.
//
// Reference: the product and SCATTER of p121.f90:89-97 (gather_scatter.f90:
// 694-850) and p123.f90:135-145 (fixed rows).
#include <algorithm>
#include <cstring>
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

// One workgroup per range of this colour (ranges[blockIdx.x]).
template <int NT>
__global__ __launch_bounds__(PF_BLOCK) void k_ebe_range(
    const float *__restrict__ Khi, const unsigned short *__restrict__ lmap, const unsigned char *__restrict__ lround,
    int nround, const int *__restrict__ rptr, const unsigned *__restrict__ ue, const int *__restrict__ ranges,
    const double *__restrict__ p, const float *__restrict__ p32, double *__restrict__ u,
    const double *__restrict__ fixstore, int fix_zero, double *__restrict__ rpart, int nels,
    const pf_state *__restrict__ st)
{
    if (skip(st, RUN_LIVE)) return;
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
        pd = p[eq];
        pl[t] = p32 ? p32[eq] : (float)pd;
        if (!(uf & UE_FIXED) && !(uf & UE_FIRST)) uold = u[eq];
        if ((uf & UE_FIXED) && (uf & UE_FIRST) && !fix_zero) fsv = fixstore[eq];
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
void launch_ranges_nt(pf_ctx *c, const float *p32)
{
    const pf_ranges &F = *c->rng;
    for (int col = 0; col < F.ncol; ++col) {
        const int nb = F.cstart[col + 1] - F.cstart[col];
        if (nb <= 0) continue;
        hipLaunchKernelGGL(k_ebe_range<NT>, dim3(nb), dim3(PF_BLOCK), 0, c->stream, c->K32hi, F.lmap, F.lround,
                           F.nround, F.rptr, F.ue, F.cperm + F.cstart[col], c->p, p32, c->u, c->fixstore,
                           c->fix_zero ? 1 : 0, F.rpart, c->nels, c->state);
    }
}

template <class T>
void free_dev(pf_ctx *c, T *&p, size_t n)
{
    pf_free(c, p, n * sizeof(T));
    p = nullptr;
}

template <class T>
int upload(pf_ctx *c, T *&dst, const std::vector<T> &src)
{
    int rc = pf_alloc(c, (void **)&dst, std::max<size_t>(src.size(), 1) * sizeof(T));
    if (rc) return rc;
    if (!src.empty())
        PF_HIP(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return PF_OK;
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

// Builds the ranges of the current steering and fixed rows on the host (g_g
// and fixstore copied back): distinct equations per range, the local index
// of every element slot, the rounds inside a range, the colours of the
// ranges and the first range of every equation.  Returns PF_ERR_ARG when the
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
    std::vector<int> gg(nslot);
    std::vector<double> fs;
    PF_HIP(hipMemcpyAsync(gg.data(), c->g_g, nslot * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (c->fixstore) {
        fs.resize((size_t)c->neq + 1);
        PF_HIP(hipMemcpyAsync(fs.data(), c->fixstore, fs.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    PF_HIP(hipStreamSynchronize(c->stream));
    std::vector<unsigned short> lmap(nslot, LM_NONE);
    std::vector<unsigned char> lround(nels, 0);
    std::vector<int> rptr(nrange + 1, 0);
    std::vector<unsigned> ue;
    ue.reserve((size_t)nslot / 2 + 16);
    std::vector<int> rcol(nrange, 0);
    std::vector<unsigned> mask((size_t)c->neq + 1, 0u);   // colours of the ranges touching each equation
    std::vector<int> tmp;
    std::vector<unsigned> rmask;
    std::vector<long> stamp;
    int nround = 1, ncol = 1;
    for (long r = 0; r < nrange; ++r) {
        const long e0 = r * R, e1 = std::min(nels, e0 + R);
        tmp.clear();
        for (long s = e0 * NT; s < e1 * NT; ++s)
            if (gg[s] > 0) tmp.push_back(gg[s]);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        rptr[r] = (int)ue.size();
        unsigned used = 0;
        for (int q : tmp) {
            ue.push_back((unsigned)q);
            used |= mask[q];
        }
        if (~used == 0u) return PF_ERR_ARG;                  // more than 32 colours
        const int col = __builtin_ctz(~used);
        rcol[r] = col;
        ncol = std::max(ncol, col + 1);
        for (int q : tmp) mask[q] |= 1u << col;
        // local indices and rounds: an element takes the lowest round no
        // earlier element of the range sharing an equation has taken
        rmask.assign(tmp.size(), 0u);
        stamp.assign(tmp.size(), -1L);
        for (long e = e0; e < e1; ++e) {
            unsigned busy = 0;
            for (int i = 0; i < NT; ++i) {
                const int q = gg[e * NT + i];
                if (q <= 0) continue;
                const int kk = (int)(std::lower_bound(tmp.begin(), tmp.end(), q) - tmp.begin());
                if (stamp[kk] == e) return PF_ERR_ARG;   // an equation twice in one element
                stamp[kk] = e;
                lmap[e * NT + i] = (unsigned short)kk;
                busy |= rmask[kk];
            }
            if (~busy == 0u) return PF_ERR_ARG;
            const int rd = __builtin_ctz(~busy);
            if (rd > 254) return PF_ERR_ARG;
            lround[e] = (unsigned char)rd;
            nround = std::max(nround, rd + 1);
            for (int i = 0; i < NT; ++i)
                if (lmap[e * NT + i] != LM_NONE) rmask[lmap[e * NT + i]] |= 1u << rd;
        }
    }
    rptr[nrange] = (int)ue.size();
    // flags: the lowest colour touching an equation writes it; fixed rows
    for (long r = 0; r < nrange; ++r)
        for (int a = rptr[r]; a < rptr[r + 1]; ++a) {
            const int q = (int)ue[a];
            if (__builtin_ctz(mask[q]) == rcol[r]) ue[a] |= UE_FIRST;
            if (!fs.empty() && fs[q] != 0.0) ue[a] |= UE_FIXED;
        }
    // ranges by colour, ascending range order inside a colour
    std::vector<int> cstart(ncol + 1, 0), cperm(nrange);
    for (long r = 0; r < nrange; ++r) cstart[rcol[r] + 1]++;
    for (int k2 = 0; k2 < ncol; ++k2) cstart[k2 + 1] += cstart[k2];
    {
        std::vector<int> fill(cstart.begin(), cstart.end() - 1);
        for (long r = 0; r < nrange; ++r) cperm[fill[rcol[r]]++] = (int)r;
    }
    c->rng = new pf_ranges();
    pf_ranges &F = *c->rng;
    F.nslot = nslot;
    F.nels = nels;
    F.nrange = nrange;
    F.nue = (long)ue.size();
    F.nround = nround;
    F.ncol = ncol;
    F.cstart = cstart;
    int rc = upload(c, F.lmap, lmap);
    if (!rc) rc = upload(c, F.lround, lround);
    if (!rc) rc = upload(c, F.rptr, rptr);
    if (!rc) rc = upload(c, F.ue, ue);
    if (!rc) rc = upload(c, F.cperm, cperm);
    if (!rc) rc = pf_alloc(c, (void **)&F.rpart, ((size_t)nrange + NF) * sizeof(double));
    if (!rc) {
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = pf_set_hip_error(e, "range upload", __FILE__, __LINE__);
    }
    if (rc) pf_ranges_drop(c);
    return rc;
}

// The mixed iteration's product, sums into u and alpha (replaces k_ebe,
// k_scatter_pu and k_fold_alpha).
int pf_ranges_iteration(pf_ctx *c, const float *p32)
{
    if (!c->rng) return PF_ERR_STATE;
    switch (c->ntot) {
    case 60: launch_ranges_nt<60>(c, p32); break;
    case 24: launch_ranges_nt<24>(c, p32); break;
    default: launch_ranges_nt<8>(c, p32); break;
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_ranges_fold(pf_ctx *c)
{
    double *fpart = c->rng->rpart + c->rng->nrange;   // NF more doubles past the range partials
    hipLaunchKernelGGL(k_fold_ranges1, dim3(NF), dim3(PF_BLOCK), 0, c->stream, c->rng->rpart,
                       (int)c->rng->nrange, fpart, c->state);
    hipLaunchKernelGGL(k_fold_alpha_ranges, dim3(1), dim3(PF_BLOCK), 0, c->stream, fpart, c->state);
    PF_HIP(hipGetLastError());
    return PF_OK;
}
