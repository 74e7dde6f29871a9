// Element-by-element PCG on the device: gather -> element products ->
// scatter, the FP64 reductions, the vector recurrences and the stopping rule.
//
// Reference flow: src/programs/5th_ed/p121/p121.f90:86-104 and
// src/programs/5th_ed/p123/p123.f90:120-151; GATHER/SCATTER/SCATTER_NOADD
// src/modules/mpi/gather_scatter.f90:547-993; DOT_PRODUCT_P/SUM_P
// src/modules/mpi/maths.f90:168-216,271-315; checon_par maths.f90:999-1069.
//
// One PCG iteration on the context's stream:
//   k_ebe         gathers p through the steering (p has a leading 0 so a
//                 restrained freedom reads 0), stages the element vector in
//                 LDS and forms u_e = K_e p_e, one 64-lane wavefront per
//                 hex20 element (lane = row, 16-byte column-pair or
//                 column-quad loads), writing utemp;
//   k_scatter_pu  sums each equation's element slots in ascending element
//                 order (the order SCATTER adds them at one rank), applies
//                 the p123 fixed-freedom rows, block sums of p.u;
//   k_fold_alpha  (one block) alpha = up / (p.u);
//   k_update      xnew = x + p*alpha, r = r - u*alpha, d = dinv*r, block
//                 sums of (r,d), (r,r), block maxima of |xnew|, |xnew-x|;
//   k_fold_update (one block) beta, the stopping rule, the iteration count
//                 and the freeze flag;
//   k_pupdate     p = d + p*beta (and, for the mixed solver, p32 = fl32(p)).
// The mixed solver adds a true-residual replacement (reliable update) on
// the iterations its fold kernel flags: k_ebe over hi+lo in FP64 with x,
// k_scatter_res (r = f - K x, d = dinv*r) and k_fold_refine.
// Every kernel returns at once when the freeze flag is set, so iterations
// enqueued after convergence are no-ops and the count equals the host loop's.
//
// FP64 vector arithmetic is compiled without contraction: every product and
// sum rounds as written; the element product is an explicit chain of fused
// multiply-adds in ascending column order.  Reductions use one fixed order
// (pf_reduce_grid in the public header) and no atomics, so results are
// reproducible bit for bit.
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "pf_internal.h"
#include "pf_device.h"

namespace {

// Element products.  GATHER: src is a neq+1 vector with src[0] = 0 and the
// element vector is src[g_g]; otherwise src is pmul (ntot,nels).  Lane i of
// an element's lane group forms row i of u_e = K_e p_e and writes it to
// out[e*NT + i], or, with rowmajor on an 8-node mesh (NT = 8), to
// out[i*nels + e].
//   KIND_F64: full K_e in FP64 column pairs, FP64 fused-multiply-add chain in
//             ascending column order -> out (double);
//   KIND_F32: packed symmetric FP32 hi, FP32 arithmetic on fl32(p) -> out
//             (float); fl32(p) is gathered from src32 when given (the
//             mixed solver keeps it beside p for hex8 meshes: half the
//             gathered bytes);
//   KIND_ACC: packed symmetric (double)hi + (double)lo, FP64 arithmetic ->
//             out (double);
//   KIND_WIDE: packed symmetric (double)hi, FP64 arithmetic on the FP64 p ->
//             out (double).
// Packed symmetric rows: lane i holds K(i, i+d mod NT) for d = 0..NT/2 and
// forms sum_d K(i,i+d) p(i+d) from p staged twice in LDS, then receives
// K(i-d,i) p(i) from lane i-d (mod NT) for d = 1..NT/2-1 by ds_bpermute,
// where lane j sends K(j,j+d) p(j): every entry is read from HBM once (the
// d = NT/2 entries, shared by rows i and i + NT/2, are stored once).  The
// products go out with non-temporal stores (they are read back by the
// scatter, not by this kernel): 2% faster FP64 products, 0.5-1% FP32.
// ebe_block forms the products of logical block lb (PF_BLOCK/64 * EPW * EPL
// consecutive elements); the kernels below choose lb.
template <int NT, int KIND, bool GATHER, typename OUT, int EPL>
__device__ __forceinline__ void ebe_block(long lb, const double *__restrict__ K64, const float *__restrict__ Khi,
                                          const float *__restrict__ Klo, const int *__restrict__ gg,
                                          const double *__restrict__ src, const float *__restrict__ src32,
                                          OUT *__restrict__ out, int nels, long kstride, int rowmajor)
{
    using L = Lanes<NT, KIND>;
    // packed layout: Q quads and R more full diagonals, then the half one (d = H)
    constexpr int D = pf_sym_diags(NT), Q = pf_sym_quads(NT), R = pf_sym_rem(NT), H = NT / 2;
    constexpr long PER = pf_sym_per(NT);
    static_assert(D == 4 * Q + R + 1, "packed layout");
    // Per wave and element slot: p_e (FP64), or p_e twice (packed kinds).
    __shared__ __attribute__((aligned(16))) double ps[PF_BLOCK / 64][EPL][KIND == KIND_F32 ? 64 : 128];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sub = lane / L::W, li = lane % L::W;
    const int i = li < NT ? li : NT - 1;
    const int gb = sub * L::W;   // first lane of the element's group
    const long ebase = (lb * (PF_BLOCK / 64) + wave) * L::EPW * EPL;
    long ev[EPL], ec[EPL];
    int g[EPL];
#pragma unroll
    for (int t = 0; t < EPL; ++t) {
        ev[t] = ebase + t * L::EPW + sub;
        ec[t] = ev[t] < nels ? ev[t] : (long)nels - 1;
        g[t] = GATHER ? __builtin_nontemporal_load(&gg[ec[t] * NT + i]) : 0;   // read once per product
    }
    if constexpr (KIND == KIND_F64) {
        // Every column pair of K_e is in flight before the dependent gather.
        d2_t kc[EPL][NT / 2];
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            const d2_t *Kp = reinterpret_cast<const d2_t *>(K64) + ec[t] * kstride * (NT * NT / 2) + i;
#pragma unroll
            for (int j = 0; j < NT / 2; ++j) kc[t][j] = __builtin_nontemporal_load(&Kp[j * NT]);
        }
#pragma unroll
        for (int t = 0; t < EPL; ++t) ps[wave][t][lane] = GATHER ? src[g[t]] : src[ec[t] * NT + i];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            const double *pw = &ps[wave][t][gb];
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < NT / 2; ++j) {
                acc = __fma_rn(kc[t][j].x, pw[2 * j], acc);
                acc = __fma_rn(kc[t][j].y, pw[2 * j + 1], acc);
            }
            if (ev[t] < nels && li < NT) {
                if (NT == 8 && rowmajor)
                    out[(long)li * nels + ev[t]] = (OUT)acc;
                else
                    __builtin_nontemporal_store((OUT)acc, &out[ev[t] * NT + li]);
            }
        }
    } else if constexpr (KIND == KIND_F32) {
        float kd[EPL][D];
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            const float *Ke = Khi + ec[t] * PER;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const f4_t v = __builtin_nontemporal_load(reinterpret_cast<const f4_t *>(Ke) + q * NT + i);
                kd[t][4 * q] = v.x; kd[t][4 * q + 1] = v.y; kd[t][4 * q + 2] = v.z; kd[t][4 * q + 3] = v.w;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) kd[t][4 * Q + r] = __builtin_nontemporal_load(Ke + 4 * Q * NT + R * i + r);
            kd[t][H] = __builtin_nontemporal_load(Ke + H * NT + i % H);
        }
        float pv[EPL];
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            if constexpr (GATHER)
                pv[t] = src32 ? src32[g[t]] : (float)src[g[t]];
            else
                pv[t] = (float)src[ec[t] * NT + i];
            float *pg = reinterpret_cast<float *>(&ps[wave][t][0]) + 2 * gb;
            if (li < NT) {
                pg[li] = pv[t];
                pg[li + NT] = pv[t];
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            const float *pg = reinterpret_cast<const float *>(&ps[wave][t][0]) + 2 * gb;
            float acc = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) acc = __fmaf_rn(kd[t][d], pg[i + d], acc);
            int j = i;   // (i - d) mod NT: the lane sending K(i-d, i) p(i)
#pragma unroll
            for (int d = 1; d < D - 1; ++d) {
                j = j == 0 ? NT - 1 : j - 1;
                const int from = gb + j;
                const float tt = __fmul_rn(kd[t][d], pv[t]);
                acc = __fadd_rn(acc, __int_as_float(__builtin_amdgcn_ds_bpermute(from << 2, __float_as_int(tt))));
            }
            if (ev[t] < nels && li < NT) {
                if (NT == 8 && rowmajor)
                    out[(long)li * nels + ev[t]] = (OUT)acc;
                else
                    __builtin_nontemporal_store((OUT)acc, &out[ev[t] * NT + li]);
            }
        }
    } else {
        constexpr bool LO = KIND == KIND_ACC;
        double kd[EPL][D];
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            const float *Kh = Khi + ec[t] * PER;
            const float *Kl = Klo + ec[t] * PER;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const f4_t h = __builtin_nontemporal_load(reinterpret_cast<const f4_t *>(Kh) + q * NT + i);
                if constexpr (LO) {
                    const f4_t l = __builtin_nontemporal_load(reinterpret_cast<const f4_t *>(Kl) + q * NT + i);
                    kd[t][4 * q] = (double)h.x + (double)l.x;
                    kd[t][4 * q + 1] = (double)h.y + (double)l.y;
                    kd[t][4 * q + 2] = (double)h.z + (double)l.z;
                    kd[t][4 * q + 3] = (double)h.w + (double)l.w;
                } else {
                    kd[t][4 * q] = (double)h.x;
                    kd[t][4 * q + 1] = (double)h.y;
                    kd[t][4 * q + 2] = (double)h.z;
                    kd[t][4 * q + 3] = (double)h.w;
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
                kd[t][4 * Q + r] = LO ? (double)__builtin_nontemporal_load(Kh + 4 * Q * NT + R * i + r) +
                                            (double)__builtin_nontemporal_load(Kl + 4 * Q * NT + R * i + r)
                                      : (double)__builtin_nontemporal_load(Kh + 4 * Q * NT + R * i + r);
            kd[t][H] = LO ? (double)__builtin_nontemporal_load(Kh + H * NT + i % H) +
                                (double)__builtin_nontemporal_load(Kl + H * NT + i % H)
                          : (double)__builtin_nontemporal_load(Kh + H * NT + i % H);
        }
        double pv[EPL];
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            pv[t] = GATHER ? src[g[t]] : src[ec[t] * NT + i];
            double *pg = &ps[wave][t][0] + 2 * gb;
            if (li < NT) {
                pg[li] = pv[t];
                pg[li + NT] = pv[t];
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int t = 0; t < EPL; ++t) {
            const double *pg = &ps[wave][t][0] + 2 * gb;
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) acc = __fma_rn(kd[t][d], pg[i + d], acc);
            int j = i;   // (i - d) mod NT
#pragma unroll
            for (int d = 1; d < D - 1; ++d) {
                j = j == 0 ? NT - 1 : j - 1;
                const int from = gb + j;
                acc = __dadd_rn(acc, bpermute64(from << 2, __dmul_rn(kd[t][d], pv[t])));
            }
            if (ev[t] < nels && li < NT) {
                if (NT == 8 && rowmajor)
                    out[(long)li * nels + ev[t]] = (OUT)acc;
                else
                    __builtin_nontemporal_store((OUT)acc, &out[ev[t] * NT + li]);
            }
        }
    }
}

// nlb logical blocks; with LOOP a grid smaller than nlb (the gated
// true-residual products) loops over them, otherwise every workgroup takes
// one.
template <int NT, int KIND, bool GATHER, typename OUT, bool LOOP = false, int EPLT = Lanes<NT, KIND>::EPL>
__global__ __launch_bounds__(PF_BLOCK) void k_ebe(const double *__restrict__ K64, const float *__restrict__ Khi,
                                                 const float *__restrict__ Klo, const int *__restrict__ gg,
                                                 const double *__restrict__ src, const float *__restrict__ src32,
                                                 OUT *__restrict__ out, int nels, long kstride,
                                                 const pf_state *__restrict__ st, int gate, long nlb, int rowmajor)
{
    if (skip(st, gate)) return;
    const long lb0 = xcd_block(blockIdx.x, gridDim.x);
    if constexpr (LOOP) {
        for (long lb = lb0; lb < nlb; lb += gridDim.x) {
            ebe_block<NT, KIND, GATHER, OUT, EPLT>(lb, K64, Khi, Klo, gg, src, src32, out, nels, kstride, rowmajor);
            __builtin_amdgcn_wave_barrier();   // this wave's LDS reads finish before the next block's writes
        }
    } else {
        ebe_block<NT, KIND, GATHER, OUT, EPLT>(lb0, K64, Khi, Klo, gg, src, src32, out, nels, kstride, rowmajor);
    }
}

// Sum of utemp over the element slots [a, b) of one equation, in slot
// (= ascending element) order.  The slot indices of up to 8 slots -- every
// equation of a hex8 or hex20 mesh of bricks has at most 8 -- are loaded
// before the dependent utemp loads; longer lists continue in a loop.
template <typename UT>
__device__ __forceinline__ double slot_sum(const int *__restrict__ slot, const UT *__restrict__ utemp, int a, int b)
{
    int sl[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) sl[j] = a + j < b ? slot[a + j] : -1;
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = sl[j] >= 0 ? (double)utemp[sl[j]] : 0.0;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (sl[j] >= 0) s = __dadd_rn(s, v[j]);
    for (int k = a + 8; k < b; ++k) s = __dadd_rn(s, (double)utemp[slot[k]]);
    return s;
}

// slot_sum with the slot indices of [a, b) read from LDS, where lds[k]
// holds slot[a0 + k].
template <typename UT>
__device__ __forceinline__ double slot_sum_lds(const int *lds, int a0, const UT *__restrict__ utemp, int a, int b)
{
    int sl[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) sl[j] = a + j < b ? lds[a - a0 + j] : -1;
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = sl[j] >= 0 ? (double)utemp[sl[j]] : 0.0;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (sl[j] >= 0) s = __dadd_rn(s, v[j]);
    for (int k = a + 8; k < b; ++k) s = __dadd_rn(s, (double)utemp[lds[k - a0]]);
    return s;
}

// SCATTER fused with p.u.  u and p have the leading 0 slot.  A fixed
// freedom's u is p*store (p123; p122's first plastic iteration) or, with
// fix_zero, 0 (p122's later plastic iterations).  Thread t of the grid takes
// equations t, t + stride, ... and adds their p.u terms in that order; each
// wavefront first copies the slot indices of its 64 equations (one
// contiguous range of the slot list) into LDS with whole-line loads, then
// every lane reads its own slots from there (more than SC_CAP of them: read
// from memory as they are).
// Block lb of G (the problem's reduction grid) does the work thread
// lb*PF_BLOCK + t of a G-block grid would do.
constexpr int SC_CAP = 1024;
template <typename UT>
__device__ __forceinline__ void scatter_pu_block(int lb, int G, const int *__restrict__ ptr,
                                                 const int *__restrict__ slot, const UT *__restrict__ utemp,
                                                 const double *__restrict__ fixstore, const double *__restrict__ p,
                                                 double *__restrict__ u, int neq, int fix_zero, double *partials)
{
    __shared__ int sls[PF_BLOCK / 64][SC_CAP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[1] = {0.0};
    const int stride = G * PF_BLOCK;
    for (int q0 = lb * PF_BLOCK + wave * 64; q0 < neq; q0 += stride) {
        const int q = q0 + lane;
        const bool live = q < neq;
        const int a = live ? __builtin_nontemporal_load(&ptr[q]) : 0, b = live ? __builtin_nontemporal_load(&ptr[q + 1]) : 0;
        const int a0 = ptr[q0], a1 = ptr[q0 + 64 < neq ? q0 + 64 : neq];
        double s = 0.0;
        if (a1 - a0 <= SC_CAP) {
            // the slot list is read once per iteration: for FP32 products,
            // streaming loads, which leave the cache to the products read
            // back here (FP64 products lost time with them)
            for (int k = lane; k < a1 - a0; k += 64)
                sls[wave][k] = sizeof(UT) == 4 ? __builtin_nontemporal_load(&slot[a0 + k]) : slot[a0 + k];
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (live) s = slot_sum_lds(&sls[wave][0], a0, utemp, a, b);
            __builtin_amdgcn_wave_barrier();   // these reads finish before the next range is written
        } else if (live) {
            s = slot_sum(slot, utemp, a, b);
        }
        if (live) {
            const double pq = p[q + 1];
            if (fixstore) {
                const double fs = fixstore[q + 1];
                if (fs != 0.0) s = fix_zero ? 0.0 : __dmul_rn(pq, fs);
            }
            u[q + 1] = s;
            acc[0] = __dadd_rn(acc[0], __dmul_rn(pq, s));
        }
    }
    constexpr int qid[1] = {PF_Q_PU};
    constexpr bool mx[1] = {false};
    block_to_partials<1>(acc, qid, mx, partials, lb);
}

template <typename UT>
__global__ __launch_bounds__(PF_BLOCK) void k_scatter_pu(const int *__restrict__ ptr, const int *__restrict__ slot,
                                                        const UT *__restrict__ utemp,
                                                        const double *__restrict__ fixstore,
                                                        const double *__restrict__ p, double *__restrict__ u,
                                                        int neq, int fix_zero, double *partials, const pf_state *st)
{
    if (skip(st, RUN_LIVE)) return;
    scatter_pu_block<UT>(blockIdx.x, gridDim.x, ptr, slot, utemp, fixstore, p, u, neq, fix_zero, partials);
}

__device__ __forceinline__ void fold_alpha_block(const double *partials, int G, pf_state *st)
{
    constexpr int qid[1] = {PF_Q_PU};
    constexpr bool mx[1] = {false};
    double out[1];
    fold_partials<1>(out, qid, mx, partials, G);
    if (threadIdx.x == 0) {
        const double alpha = st->up / out[0];
        st->alpha = alpha;
        if (!isfinite(alpha)) {
            st->breakdown = 1;
            st->frozen = 1;
        }
    }
}

__global__ __launch_bounds__(PF_BLOCK) void k_fold_alpha(const double *partials, int G, pf_state *st)
{
    if (skip(st, RUN_LIVE)) return;
    fold_alpha_block(partials, G, st);
}

// x/r/d updates and the partial (r,d) with, for checon_par, max|xnew| and
// max|xnew-x|, or, for the residual rule, (r,r).
template <bool CHECON>
__device__ __forceinline__ void update_block(int lb, int G, double *__restrict__ x, double *__restrict__ r,
                                             double *__restrict__ d, const double *__restrict__ p,
                                             const double *__restrict__ u, const double *__restrict__ dinv, int neq,
                                             double *partials, const pf_state *st)
{
    const double alpha = st->alpha;
    constexpr int NQ = CHECON ? 3 : 2;
    double v[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = 0.0;
    const int stride = G * PF_BLOCK;
    for (int q = lb * PF_BLOCK + threadIdx.x + 1; q <= neq; q += stride) {
        // x, r, u and dinv are next touched a whole iteration later
        // (streaming loads and stores); d is read again by the p update
        // right after
        const double xo = __builtin_nontemporal_load(&x[q]);
        const double xn = __dadd_rn(xo, __dmul_rn(p[q], alpha));
        const double rq = __dsub_rn(__builtin_nontemporal_load(&r[q]),
                                    __dmul_rn(__builtin_nontemporal_load(&u[q]), alpha));
        const double dq = __dmul_rn(__builtin_nontemporal_load(&dinv[q]), rq);
        __builtin_nontemporal_store(xn, &x[q]);
        __builtin_nontemporal_store(rq, &r[q]);
        d[q] = dq;
        v[0] = __dadd_rn(v[0], __dmul_rn(rq, dq));
        if constexpr (CHECON) {
            v[1] = fmax(v[1], fabs(xn));
            v[2] = fmax(v[2], fabs(__dsub_rn(xn, xo)));
        } else {
            v[1] = __dadd_rn(v[1], __dmul_rn(rq, rq));
        }
    }
    if constexpr (CHECON) {
        constexpr int qid[3] = {PF_Q_RD, PF_Q_MAXX, PF_Q_MAXDX};
        constexpr bool mx[3] = {false, true, true};
        block_to_partials<3>(v, qid, mx, partials, lb);
    } else {
        constexpr int qid[2] = {PF_Q_RD, PF_Q_RR};
        constexpr bool mx[2] = {false, false};
        block_to_partials<2>(v, qid, mx, partials, lb);
    }
}

template <bool CHECON>
__global__ __launch_bounds__(PF_BLOCK) void k_update(double *__restrict__ x, double *__restrict__ r,
                                                    double *__restrict__ d, const double *__restrict__ p,
                                                    const double *__restrict__ u, const double *__restrict__ dinv,
                                                    int neq, double *partials, const pf_state *st)
{
    if (skip(st, RUN_LIVE)) return;
    update_block<CHECON>(blockIdx.x, gridDim.x, x, r, d, p, u, dinv, neq, partials, st);
}

// End of an iteration: beta = rd/up, the stopping rule, the count.
__device__ void finish_iteration(pf_state *st, double rd, double rr, double maxx, double maxdx)
{
    const double beta = rd / st->up;
    const double crit = st->stop_mode == PF_STOP_CHECON ? maxdx / maxx : sqrt(rr / st->rr0);
    st->beta = beta;
    st->up = rd;
    st->rr = rr;
    st->crit = crit;
    st->iters += 1;
    st->pfresh = 0;
    st->converged = crit <= st->tol;
    if (!isfinite(beta) || !isfinite(rd)) st->breakdown = 1;
    if (st->converged || st->iters >= st->limit || st->breakdown) st->frozen = 1;
}

template <bool CHECON>
__device__ __forceinline__ void fold_update_block(const double *partials, int G, pf_state *st)
{
    double out[4] = {0.0, 0.0, 0.0, 0.0};   // rd, rr, max|xnew|, max|xnew-x|
    if constexpr (CHECON) {
        constexpr int qid[3] = {PF_Q_RD, PF_Q_MAXX, PF_Q_MAXDX};
        constexpr bool mx[3] = {false, true, true};
        double o[3];
        fold_partials<3>(o, qid, mx, partials, G);
        out[0] = o[0];
        out[2] = o[1];
        out[3] = o[2];
    } else {
        constexpr int qid[2] = {PF_Q_RD, PF_Q_RR};
        constexpr bool mx[2] = {false, false};
        double o[2];
        fold_partials<2>(o, qid, mx, partials, G);
        out[0] = o[0];
        out[1] = o[1];
    }
    if (threadIdx.x == 0) {
        if (st->precision == PF_PREC_MIXED || st->precision == PF_PREC_MIXED64) {
            // Reliable update: replace the recurrence residual by the true
            // one when it has dropped by delta since the last replacement,
            // periodically if asked, and before accepting convergence.
            st->since_refine += 1;
            const double rr = out[1];
            const bool trig = rr <= st->delta2 * st->rr_ref || sqrt(rr / st->rr0) <= st->tol ||
                              (st->refine_every > 0 && st->since_refine >= st->refine_every);
            if (trig && isfinite(rr)) {
                st->refine = 1;
                return;   // finish_iteration runs after the replacement
            }
        }
        finish_iteration(st, out[0], out[1], out[2], out[3]);
    }
}

template <bool CHECON>
__global__ __launch_bounds__(PF_BLOCK) void k_fold_update(const double *partials, int G, pf_state *st)
{
    if (skip(st, RUN_LIVE)) return;
    fold_update_block<CHECON>(partials, G, st);
}

// True residual r = f - K x (K from hi + lo, or the fixed-row override
// x*store), d = dinv*r, partial (r,d), (r,r).
__device__ __forceinline__ void scatter_res_block(int lb, int G, const int *__restrict__ ptr,
                                                  const int *__restrict__ slot, const double *__restrict__ utemp,
                                                  const double *__restrict__ fixstore, const double *__restrict__ x,
                                                  const double *__restrict__ f, double *__restrict__ r,
                                                  double *__restrict__ d, const double *__restrict__ dinv, int neq,
                                                  double *partials)
{
    double v[2] = {0.0, 0.0};
    const int stride = G * PF_BLOCK;
    for (int q = lb * PF_BLOCK + threadIdx.x; q < neq; q += stride) {
        double s = slot_sum(slot, utemp, ptr[q], ptr[q + 1]);
        if (fixstore) {
            const double fs = fixstore[q + 1];
            if (fs != 0.0) s = __dmul_rn(x[q + 1], fs);
        }
        const double rq = __dsub_rn(f[q + 1], s);
        const double dq = __dmul_rn(dinv[q + 1], rq);
        r[q + 1] = rq;
        d[q + 1] = dq;
        v[0] = __dadd_rn(v[0], __dmul_rn(rq, dq));
        v[1] = __dadd_rn(v[1], __dmul_rn(rq, rq));
    }
    constexpr int qid[2] = {PF_Q_RD, PF_Q_RR};
    constexpr bool mx[2] = {false, false};
    block_to_partials<2>(v, qid, mx, partials, lb);
}

__global__ __launch_bounds__(PF_BLOCK) void k_scatter_res(const int *__restrict__ ptr, const int *__restrict__ slot,
                                                         const double *__restrict__ utemp,
                                                         const double *__restrict__ fixstore,
                                                         const double *__restrict__ x, const double *__restrict__ f,
                                                         double *__restrict__ r, double *__restrict__ d,
                                                         const double *__restrict__ dinv, int neq, double *partials,
                                                         const pf_state *st)
{
    if (skip(st, RUN_REFINE)) return;
    scatter_res_block(blockIdx.x, gridDim.x, ptr, slot, utemp, fixstore, x, f, r, d, dinv, neq, partials);
}

__device__ __forceinline__ void fold_refine_block(const double *partials, int G, pf_state *st)
{
    constexpr int qid[2] = {PF_Q_RD, PF_Q_RR};
    constexpr bool mx[2] = {false, false};
    double out[2];
    fold_partials<2>(out, qid, mx, partials, G);
    if (threadIdx.x == 0) {
        st->refine = 0;
        st->refinements += 1;
        st->since_refine = 0;
        st->rr_ref = out[1];
        finish_iteration(st, out[0], out[1], 0.0, 0.0);
    }
}

__global__ __launch_bounds__(PF_BLOCK) void k_fold_refine(const double *partials, int G, pf_state *st)
{
    if (skip(st, RUN_REFINE)) return;
    fold_refine_block(partials, G, st);
}

// p = d + p*beta, and p32 = fl32(p) when the FP32 products gather it.
__device__ __forceinline__ void pupdate_block(int lb, int G, double *__restrict__ p, float *__restrict__ p32,
                                              const double *__restrict__ d, int neq, const pf_state *__restrict__ st)
{
    const double beta = st->beta;
    const int stride = G * PF_BLOCK;
    for (int q = lb * PF_BLOCK + threadIdx.x + 1; q <= neq; q += stride) {
        // streaming loads: d and the old p are not read again this iteration
        // (measured: this pass slows by 7 us, the element products that follow
        // gain 10-70 us)
        const double pq = __dadd_rn(__builtin_nontemporal_load(&d[q]), __dmul_rn(__builtin_nontemporal_load(&p[q]), beta));
        __builtin_nontemporal_store(pq, &p[q]);   // read next by the gathers, not from this L2 line
        if (p32) __builtin_nontemporal_store((float)pq, &p32[q]);
    }
}

__global__ __launch_bounds__(PF_BLOCK) void k_pupdate(double *__restrict__ p, float *__restrict__ p32,
                                                     const double *__restrict__ d, int neq,
                                                     const pf_state *__restrict__ st)
{
    if (skip(st, RUN_LIVE)) return;
    pupdate_block(blockIdx.x, gridDim.x, p, p32, d, neq, st);
}

// PCG start: dinv = 1/diag, d = dinv*r, p = d, x = x0 (or 0), f = r; partial
// (r,d), (r,r).
__global__ __launch_bounds__(PF_BLOCK) void k_pcg_init(double *__restrict__ x, const double *__restrict__ r,
                                                      double *__restrict__ d, double *__restrict__ p,
                                                      float *__restrict__ p32,
                                                      double *__restrict__ dinv, const double *__restrict__ diag,
                                                      double *__restrict__ f, const double *__restrict__ x0, int neq,
                                                      int keep_dinv, double *partials)
{
    double v[2] = {0.0, 0.0};
    const int stride = gridDim.x * PF_BLOCK;
    for (int q = blockIdx.x * PF_BLOCK + threadIdx.x + 1; q <= neq; q += stride) {
        const double di = keep_dinv ? dinv[q] : 1.0 / diag[q];
        const double rq = r[q];
        const double dq = __dmul_rn(di, rq);
        dinv[q] = di;
        d[q] = dq;
        p[q] = dq;
        if (p32) p32[q] = (float)dq;
        x[q] = x0 ? x0[q - 1] : 0.0;
        f[q] = rq;
        v[0] = __dadd_rn(v[0], __dmul_rn(rq, dq));
        v[1] = __dadd_rn(v[1], __dmul_rn(rq, rq));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        x[0] = 0.0; d[0] = 0.0; p[0] = 0.0; dinv[0] = 0.0; f[0] = 0.0;
        if (p32) p32[0] = 0.0f;
    }
    constexpr int qid[2] = {PF_Q_RD, PF_Q_RR};
    constexpr bool mx[2] = {false, false};
    block_to_partials<2>(v, qid, mx, partials, blockIdx.x);
}

__global__ __launch_bounds__(PF_BLOCK) void k_fold_init(const double *partials, int G, pf_state *st)
{
    constexpr int qid[2] = {PF_Q_RD, PF_Q_RR};
    constexpr bool mx[2] = {false, false};
    double out[2];
    fold_partials<2>(out, qid, mx, partials, G);
    if (threadIdx.x == 0) {
        st->up = out[0];
        st->rr0 = out[1];
        st->rr = out[1];
        st->rr_ref = out[1];
        st->iters = 0;
        st->breakdown = !isfinite(out[0]) || !isfinite(out[1]);
        // A zero right-hand side is already solved by x = 0.
        st->converged = out[1] == 0.0;
        st->frozen = st->converged || st->breakdown;
        st->crit = 0.0;
    }
}

// ------------------------------------------- reference-order reductions
// DOT_PRODUCT at one rank (maths.f90:168-216): s = 0; s = s + a(i)*b(i) for
// i = 1..n in ascending order, each product and sum rounded as written (b
// null: SUM, s = s + a(i), maths.f90:271-315).  The chain is sequential by
// definition, so one wavefront walks it: the lanes load and multiply 64
// consecutive terms, a uniform register takes them in lane order through
// v_readlane, and the next 256 terms are loaded while the current ones are
// added.
__device__ __forceinline__ double readlane64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ double seq_sum(const double *__restrict__ a, const double *__restrict__ b, long n)
{
    constexpr int U = 4;
    const int lane = threadIdx.x & 63;
    double t[U], nx[U];
    auto load = [&](long i0, double (&v)[U]) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const long i = i0 + 64 * k + lane;
            v[k] = i < n ? (b ? __dmul_rn(a[i], b[i]) : a[i]) : 0.0;
        }
    };
    double s = 0.0;
    load(0, t);
    for (long i0 = 0; i0 < n; i0 += 64 * U) {
        load(i0 + 64 * U, nx);
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const long m = n - (i0 + 64 * k);
            if (m >= 64) {
#pragma unroll
                for (int l = 0; l < 64; ++l) s = __dadd_rn(s, readlane64(t[k], l));
            } else {
                for (int l = 0; l < (int)m; ++l) s = __dadd_rn(s, readlane64(t[k], l));
            }
        }
#pragma unroll
        for (int k = 0; k < U; ++k) t[k] = nx[k];
    }
    return s;
}

// Up to two reference-order sums (a[k], b[k] -> quantity qs[k]) and up to
// two maxima (the G block maxima of quantity qm[k], exact in any order),
// written to slot 0 of their quantities: the fold kernels then run with one
// block partial.  One wavefront.
struct pf_seq_args {
    const double *a[2];
    const double *b[2];
    int qs[2];
    int ns;
    int qm[2];
    int nm;
};

__global__ __launch_bounds__(64) void k_seq_reduce(pf_seq_args A, long n, int G, double *partials,
                                                  const pf_state *st, int gate)
{
    if (skip(st, gate)) return;
    for (int k = 0; k < A.nm; ++k) {
        const double *P = partials + (long)A.qm[k] * MAXG;
        double m = 0.0;
        for (int j = threadIdx.x; j < G; j += 64) m = fmax(m, P[j]);
        m = wave_max(m);
        __builtin_amdgcn_wave_barrier();   // every lane has read P[0] before lane 0 writes it
        if (threadIdx.x == 0) partials[(long)A.qm[k] * MAXG] = m;
    }
    for (int k = 0; k < A.ns; ++k) {
        const double s = seq_sum(A.a[k], A.b[k], n);
        if (threadIdx.x == 0) partials[(long)A.qs[k] * MAXG] = s;
    }
}

// ---------------------------------------------------------- primitives
typedef unsigned u4_t __attribute__((ext_vector_type(4)));

// Device copy in 16-byte words: each workgroup copies contiguous chunks of
// COPY_CHUNK words (256 KiB), four loads in flight per lane before their
// stores, then the chunk's remainder one word at a time.  Contiguous chunks
// per workgroup measured 5.3 TB/s over 1 GiB against 4.7 for a grid-stride
// walk with the same loads in flight.
constexpr long COPY_CHUNK = 16384;

__global__ __launch_bounds__(256) void k_copy(const u4_t *__restrict__ src, u4_t *__restrict__ dst, long n16)
{
    for (long base = (long)blockIdx.x * COPY_CHUNK; base < n16; base += (long)gridDim.x * COPY_CHUNK) {
        const long end = base + COPY_CHUNK < n16 ? base + COPY_CHUNK : n16;
        long i = base + threadIdx.x;
        for (; i + 768 < end; i += 1024) {
            const u4_t v0 = src[i], v1 = src[i + 256], v2 = src[i + 512], v3 = src[i + 768];
            dst[i] = v0;
            dst[i + 256] = v1;
            dst[i + 512] = v2;
            dst[i + 768] = v3;
        }
        for (; i < end; i += 256) dst[i] = src[i];
    }
}

// Whole-array vector statements: 0 fill, 1 w = x + y*a, 2 z = x*y, 3 y = 1/x.
__global__ __launch_bounds__(256) void k_vec(int op, double *w, const double *x, const double *y, double a, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        switch (op) {
        case 0: w[i] = a; break;
        case 1: w[i] = __dadd_rn(x[i], __dmul_rn(y[i], a)); break;
        case 2: w[i] = __dmul_rn(x[i], y[i]); break;
        default: w[i] = 1.0 / x[i]; break;
        }
    }
}

__global__ void k_gather(const int *__restrict__ gg, const double *__restrict__ p, double *__restrict__ pmul, long n)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const int g = gg[k];
        pmul[k] = g > 0 ? p[g - 1] : 0.0;
    }
}

__global__ void k_scatter_prim(const int *__restrict__ ptr, const int *__restrict__ slot,
                               const double *__restrict__ utemp, double *__restrict__ u, int neq, int add)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < neq; q += gridDim.x * blockDim.x) {
        const int a = ptr[q], b = ptr[q + 1];
        if (add) {
            double s = 0.0;
            for (int k = a; k < b; ++k) s = __dadd_rn(s, utemp[slot[k]]);
            u[q] = __dadd_rn(u[q], s);
        } else if (b > a) {
            u[q] = utemp[slot[b - 1]];   // the last (highest) element wins
        }
    }
}

// Reductions for the primitives: mode 0 = sum a*b, 1 = sum a,
// 2 = max|a| and max|a-b| then b = a (checon_par).
__global__ __launch_bounds__(PF_BLOCK) void k_reduce_prim(const double *__restrict__ a, double *__restrict__ b,
                                                         long n, int mode, double *partials)
{
    double v[2] = {0.0, 0.0};
    const long stride = (long)gridDim.x * PF_BLOCK;
    long i = (long)blockIdx.x * PF_BLOCK + threadIdx.x;
    if (mode < 2) {
        // Four strided entries in flight per lane, summed in the same order
        // as one at a time.
        for (; i + 3 * stride < n; i += 4 * stride) {
            double x[4], y[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                x[k] = a[i + k * stride];
                y[k] = mode == 0 ? b[i + k * stride] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) v[0] = __dadd_rn(v[0], mode == 0 ? __dmul_rn(x[k], y[k]) : x[k]);
        }
    }
    for (; i < n; i += stride) {
        const double ai = a[i];
        if (mode == 0) {
            v[0] = __dadd_rn(v[0], __dmul_rn(ai, b[i]));
        } else if (mode == 1) {
            v[0] = __dadd_rn(v[0], ai);
        } else {
            v[0] = fmax(v[0], fabs(ai));
            v[1] = fmax(v[1], fabs(__dsub_rn(ai, b[i])));
            b[i] = ai;
        }
    }
    constexpr int qid[2] = {0, 1};
    if (mode == 2) {
        constexpr bool mx[2] = {true, true};
        block_to_partials<2>(v, qid, mx, partials, blockIdx.x);
    } else {
        constexpr bool mx[2] = {false, false};
        block_to_partials<2>(v, qid, mx, partials, blockIdx.x);
    }
}

__global__ __launch_bounds__(PF_BLOCK) void k_fold_prim(const double *partials, int G, int mode, double *out)
{
    constexpr int qid[2] = {0, 1};
    double o[2];
    if (mode == 2) {
        constexpr bool mx[2] = {true, true};
        fold_partials<2>(o, qid, mx, partials, G);
    } else {
        constexpr bool mx[2] = {false, false};
        fold_partials<2>(o, qid, mx, partials, G);
    }
    if (threadIdx.x == 0) {
        out[0] = o[0];
        out[1] = o[1];
    }
}

// y = K x with the fixed rows (pf_apply): scatter into y (no leading slot).
__global__ void k_scatter_fix(const int *__restrict__ ptr, const int *__restrict__ slot,
                              const double *__restrict__ utemp, const double *__restrict__ fixstore,
                              const double *__restrict__ xext, double *__restrict__ y, int neq)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < neq; q += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = ptr[q]; k < ptr[q + 1]; ++k) s = __dadd_rn(s, utemp[slot[k]]);
        if (fixstore && fixstore[q + 1] != 0.0) s = __dmul_rn(xext[q + 1], fixstore[q + 1]);
        y[q] = s;
    }
}

template <int NT, int KIND, bool GATHER, typename OUT, int EPL>
void launch_ebe_epl(pf_ctx *c, const double *src, OUT *out, const pf_state *st, int gate, const double *K64,
                    const float *src32, int rowmajor)
{
    constexpr int EPB = Lanes<NT, KIND>::EPW * EPL * (PF_BLOCK / 64);
    const long nlb = (c->nels + EPB - 1) / EPB;
    hipLaunchKernelGGL((k_ebe<NT, KIND, GATHER, OUT, false, EPL>), dim3((int)nlb), dim3(PF_BLOCK), 0, c->stream,
                       K64, c->K32hi, c->K32lo, c->g_g, src, src32, out, c->nels, pf_k64_estride(c), st, gate,
                       nlb, rowmajor);
}

template <int NT, int KIND, bool GATHER, typename OUT>
int launch_ebe_nt(pf_ctx *c, const double *src, OUT *out, const pf_state *st, int gate, const double *Kalt,
                  const float *src32, int rowmajor)
{
    const double *K64 = Kalt ? Kalt : c->K64;
    if (gate == RUN_REFINE) {
        // usually does nothing: dispatch few workgroups that loop
        const long nlb = (c->nels + Lanes<NT, KIND>::EPB - 1) / Lanes<NT, KIND>::EPB;
        const int grid = (int)std::min<long>(nlb, 2048);
        hipLaunchKernelGGL((k_ebe<NT, KIND, GATHER, OUT, true>), dim3(grid), dim3(PF_BLOCK), 0, c->stream, K64,
                           c->K32hi, c->K32lo, c->g_g, src, src32, out, c->nels, pf_k64_estride(c), st, gate,
                           nlb, rowmajor);
    } else {
        // the tuned shape of the iteration's product (an out-of-range value,
        // e.g. from before the mesh changed, keeps the built-in one)
        constexpr int E0 = Lanes<NT, KIND>::EPL;
        const int e = GATHER && c->tune_epl <= pf_epl_max(NT) ? c->tune_epl : 0;
        if constexpr (GATHER) {
            if (e == 1) launch_ebe_epl<NT, KIND, GATHER, OUT, 1>(c, src, out, st, gate, K64, src32, rowmajor);
            else if (e == 2) launch_ebe_epl<NT, KIND, GATHER, OUT, 2>(c, src, out, st, gate, K64, src32, rowmajor);
            else if (e == 4) launch_ebe_epl<NT, KIND, GATHER, OUT, (NT <= 24 ? 4 : E0)>(c, src, out, st, gate, K64, src32, rowmajor);
            else if (e == 8) launch_ebe_epl<NT, KIND, GATHER, OUT, (NT <= 8 ? 8 : E0)>(c, src, out, st, gate, K64, src32, rowmajor);
            else launch_ebe_epl<NT, KIND, GATHER, OUT, E0>(c, src, out, st, gate, K64, src32, rowmajor);
        } else {
            launch_ebe_epl<NT, KIND, GATHER, OUT, E0>(c, src, out, st, gate, K64, src32, rowmajor);
        }
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}

template <int KIND, bool GATHER, typename OUT>
int launch_ebe(pf_ctx *c, const double *src, OUT *out, const pf_state *st, int gate, const double *Kalt = nullptr,
               const float *src32 = nullptr, int rowmajor = 0)
{
    switch (c->ntot) {
    case 60: return launch_ebe_nt<60, KIND, GATHER, OUT>(c, src, out, st, gate, Kalt, src32, rowmajor);
    case 24: return launch_ebe_nt<24, KIND, GATHER, OUT>(c, src, out, st, gate, Kalt, src32, rowmajor);
    case 8: return launch_ebe_nt<8, KIND, GATHER, OUT>(c, src, out, st, gate, Kalt, src32, rowmajor);
    default: return PF_ERR_ARG;
    }
}

// FP64 element products of the stored matrices (K64, or hi+lo for FP32
// storage), without state gating: the drop-in primitive and pf_apply.
template <bool GATHER>
int launch_ebe_exact(pf_ctx *c, const double *src, double *out)
{
    if (c->K64) return launch_ebe<KIND_F64, GATHER, double>(c, src, out, nullptr, RUN_ALWAYS);
    return launch_ebe<KIND_ACC, GATHER, double>(c, src, out, nullptr, RUN_ALWAYS);
}

// With the PF_DOT_SEQ order: the reference-order sums (and the folded
// maxima) of A replace the block partials' slot 0; returns the block count
// the following fold kernel reads (1).  With the tree order: nothing, G.
int seq_reduce(pf_ctx *c, int G, const double *a0, const double *b0, int q0, const double *a1,
               const double *b1, int q1, bool maxima, long n, const pf_state *st, int gate)
{
    if (c->dot_order != PF_DOT_SEQ) return G;
    pf_seq_args A{};
    A.a[0] = a0; A.b[0] = b0; A.qs[0] = q0;
    A.a[1] = a1; A.b[1] = b1; A.qs[1] = q1;
    A.ns = a1 ? 2 : 1;
    A.qm[0] = PF_Q_MAXX; A.qm[1] = PF_Q_MAXDX;
    A.nm = maxima ? 2 : 0;
    hipLaunchKernelGGL(k_seq_reduce, dim3(1), dim3(64), 0, c->stream, A, n, G, c->partials, st, gate);
    return 1;
}

int reduce_prim(pf_ctx *c, const double *a, double *b, int64_t n, int mode, double *out2)
{
    if (n <= 0) {
        out2[0] = out2[1] = 0.0;
        return PF_OK;
    }
    int G = pf_grid_for(n);
    double *dout = c->partials + (size_t)PF_NQ * MAXG;   // spare slots past the PCG partials
    if (mode < 2 && c->dot_order == PF_DOT_SEQ)
        G = seq_reduce(c, G, a, mode == 0 ? b : nullptr, 0, nullptr, nullptr, 0, false, (long)n, nullptr,
                       RUN_ALWAYS);
    else
        hipLaunchKernelGGL(k_reduce_prim, dim3(G), dim3(PF_BLOCK), 0, c->stream, a, b, (long)n, mode,
                           c->partials);
    hipLaunchKernelGGL(k_fold_prim, dim3(1), dim3(PF_BLOCK), 0, c->stream, c->partials, G, mode, dout);
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(c->scalar_host, dout, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    out2[0] = c->scalar_host[0];
    out2[1] = c->scalar_host[1];
    return PF_OK;
}

// ------------------------------------------------- batched iterations
// Several solves advanced by one launch per kernel (pf_batch_run).  Each
// solve keeps its own buffers and state; a launch's blocks are dealt to the
// solves in ranges (est: first logical element block of each solve in the
// element launches; qst: first block of each solve in the equation
// launches, G blocks per solve as its own launches use), and every block
// runs the same body with the same local block index and grid size as the
// solve's own launch would: the results are those of the solve alone, bit
// for bit.
struct pf_bslot {
    const double *K64;
    const float *Khi, *Klo;
    const int *g_g, *csr_ptr, *slot;   // slot: the list matching the products' layout
    double *x, *r, *p, *d, *u, *dinv, *f, *utemp;
    float *p32, *utemp32;
    const double *fixstore;
    double *partials;
    pf_state *st;
    long kstride;
    int nels, neq, G, fix_zero, rowmajor;
    int solo;   // runs its own launches inside the batch's run: no blocks here
};

// A pointer read from memory (the table) is generic to the compiler, which
// then emits flat loads and stores -- counted against the LDS wait counter
// too, so every LDS exchange of the element product would also wait for the
// matrix loads in flight.  Re-forming each one as a global-address-space
// pointer (through its integer value: a direct round trip of casts folds
// back to the generic pointer) lets them be global loads and stores, as with
// kernel arguments.
template <typename P>
__device__ __forceinline__ P *as_global(P *p)
{
    return (P *)(__attribute__((address_space(1))) P *)(unsigned long)p;
}

__device__ __forceinline__ pf_bslot slot_of(const pf_bslot *__restrict__ T, int s)
{
    const pf_bslot &R = T[s];
    pf_bslot S = R;
    S.K64 = as_global(R.K64); S.Khi = as_global(R.Khi); S.Klo = as_global(R.Klo);
    S.g_g = as_global(R.g_g); S.csr_ptr = as_global(R.csr_ptr); S.slot = as_global(R.slot);
    S.x = as_global(R.x); S.r = as_global(R.r); S.p = as_global(R.p); S.d = as_global(R.d);
    S.u = as_global(R.u); S.dinv = as_global(R.dinv); S.f = as_global(R.f); S.utemp = as_global(R.utemp);
    S.p32 = as_global(R.p32); S.utemp32 = as_global(R.utemp32); S.fixstore = as_global(R.fixstore);
    S.partials = as_global(R.partials); S.st = as_global(R.st);
    return S;
}

// The solve owning block b: the last solve whose first block is <= b (at
// most 64 solves, one lane each).
template <typename I>
__device__ __forceinline__ int bslot_find(const I *__restrict__ start, int ns, long b)
{
    const int lane = threadIdx.x & 63;
    const bool le = lane < ns && (long)start[lane] <= b;
    return __popcll(__ballot(le)) - 1;
}

// The work of block b of each batched launch (the bodies below are shared by
// the launches per kernel and by the persistent kernel).
// Element products of every solve: the iteration's product of p (gate
// RUN_LIVE) or, with REF, the true-residual product of x (gate RUN_REFINE).
template <int NT, int KIND, typename OUT, bool REF>
__device__ __forceinline__ void bs_ebe(const pf_bslot *__restrict__ T, const long *__restrict__ est, int ns, long gb,
                                       int gate)
{
    constexpr int EPL = Lanes<NT, KIND>::EPL;
    const int s = bslot_find(est, ns, gb);
    const pf_bslot S = slot_of(T, s);
    if (skip(S.st, gate)) return;
    const long lb = gb - est[s];
    if constexpr (REF)
        ebe_block<NT, KIND, true, OUT, EPL>(lb, S.K64, S.Khi, S.Klo, S.g_g, S.x, nullptr, (OUT *)S.utemp, S.nels,
                                            S.kstride, S.rowmajor);
    else if constexpr (KIND == KIND_F32)
        ebe_block<NT, KIND, true, OUT, EPL>(lb, S.K64, S.Khi, S.Klo, S.g_g, S.p, S.p32, (OUT *)S.utemp32, S.nels,
                                            S.kstride, S.rowmajor);
    else
        ebe_block<NT, KIND, true, OUT, EPL>(lb, S.K64, S.Khi, S.Klo, S.g_g, S.p, nullptr, (OUT *)S.utemp, S.nels,
                                            S.kstride, S.rowmajor);
}

// The true-residual products of the solves due for a replacement, solve by
// solve (a solve not due costs one state read per workgroup), blocks
// b0, b0 + nb, ... of each.
template <int NT>
__device__ __forceinline__ void bs_ebe_refine(const pf_bslot *__restrict__ T, const long *__restrict__ est, int ns,
                                              long total, long b0, long nb)
{
    for (int s = 0; s < ns; ++s) {
        if (skip(as_global(T[s].st), RUN_REFINE)) continue;
        const long end = s + 1 < ns ? est[s + 1] : total;
        for (long gb = est[s] + b0; gb < end; gb += nb) {
            bs_ebe<NT, KIND_ACC, double, true>(T, est, ns, gb, RUN_REFINE);
            __builtin_amdgcn_wave_barrier();   // this wave's LDS reads finish before the next block's writes
        }
    }
}

template <typename UT>
__device__ __forceinline__ void bs_scatter_pu(const pf_bslot *__restrict__ T, const int *__restrict__ qst, int ns,
                                              int b)
{
    const int s = bslot_find(qst, ns, b);
    const pf_bslot S = slot_of(T, s);
    if (skip(S.st, RUN_LIVE)) return;
    const UT *ut;
    if constexpr (sizeof(UT) == 4) ut = (const UT *)S.utemp32;
    else ut = (const UT *)S.utemp;
    scatter_pu_block<UT>(b - qst[s], S.G, S.csr_ptr, S.slot, ut, S.fixstore, S.p, S.u, S.neq, S.fix_zero,
                         S.partials);
}

template <bool CHECON>
__device__ __forceinline__ void bs_update(const pf_bslot *__restrict__ T, const int *__restrict__ qst, int ns, int b)
{
    const int s = bslot_find(qst, ns, b);
    const pf_bslot S = slot_of(T, s);
    if (skip(S.st, RUN_LIVE)) return;
    update_block<CHECON>(b - qst[s], S.G, S.x, S.r, S.d, S.p, S.u, S.dinv, S.neq, S.partials, S.st);
}

__device__ __forceinline__ void bs_scatter_res(const pf_bslot *__restrict__ T, const int *__restrict__ qst, int ns,
                                               int b)
{
    const int s = bslot_find(qst, ns, b);
    const pf_bslot S = slot_of(T, s);
    if (skip(S.st, RUN_REFINE)) return;
    scatter_res_block(b - qst[s], S.G, S.csr_ptr, S.slot, S.utemp, S.fixstore, S.x, S.f, S.r, S.d, S.dinv, S.neq,
                      S.partials);
}

__device__ __forceinline__ void bs_pupdate(const pf_bslot *__restrict__ T, const int *__restrict__ qst, int ns, int b)
{
    const int s = bslot_find(qst, ns, b);
    const pf_bslot S = slot_of(T, s);
    if (skip(S.st, RUN_LIVE)) return;
    pupdate_block(b - qst[s], S.G, S.p, S.p32, S.d, S.neq, S.st);
}

// The folds of solve s: 0 alpha, 1 update (checon), 2 update (residual),
// 3 true-residual replacement.
template <int WHICH>
__device__ __forceinline__ void bs_fold(const pf_bslot *__restrict__ T, int s)
{
    const pf_bslot S = slot_of(T, s);
    if (S.solo || skip(S.st, WHICH == 3 ? RUN_REFINE : RUN_LIVE)) return;
    if constexpr (WHICH == 0) fold_alpha_block(S.partials, S.G, S.st);
    else if constexpr (WHICH == 1) fold_update_block<true>(S.partials, S.G, S.st);
    else if constexpr (WHICH == 2) fold_update_block<false>(S.partials, S.G, S.st);
    else fold_refine_block(S.partials, S.G, S.st);
}

// One launch per kernel of the iteration.
template <int NT, int KIND, typename OUT>
__global__ __launch_bounds__(PF_BLOCK) void kb_ebe(const pf_bslot *__restrict__ T, const long *__restrict__ est, int ns)
{
    bs_ebe<NT, KIND, OUT, false>(T, est, ns, xcd_block(blockIdx.x, gridDim.x), RUN_LIVE);
}

template <int NT>
__global__ __launch_bounds__(PF_BLOCK) void kb_ebe_refine(const pf_bslot *__restrict__ T, const long *__restrict__ est,
                                                         int ns, long total)
{
    bs_ebe_refine<NT>(T, est, ns, total, blockIdx.x, gridDim.x);
}

template <typename UT>
__global__ __launch_bounds__(PF_BLOCK) void kb_scatter_pu(const pf_bslot *__restrict__ T, const int *__restrict__ qst,
                                                         int ns)
{
    bs_scatter_pu<UT>(T, qst, ns, blockIdx.x);
}

template <bool CHECON>
__global__ __launch_bounds__(PF_BLOCK) void kb_update(const pf_bslot *__restrict__ T, const int *__restrict__ qst,
                                                     int ns)
{
    bs_update<CHECON>(T, qst, ns, blockIdx.x);
}

__global__ __launch_bounds__(PF_BLOCK) void kb_scatter_res(const pf_bslot *__restrict__ T, const int *__restrict__ qst,
                                                          int ns)
{
    bs_scatter_res(T, qst, ns, blockIdx.x);
}

__global__ __launch_bounds__(PF_BLOCK) void kb_pupdate(const pf_bslot *__restrict__ T, const int *__restrict__ qst,
                                                      int ns)
{
    bs_pupdate(T, qst, ns, blockIdx.x);
}

template <int WHICH>
__global__ __launch_bounds__(PF_BLOCK) void kb_fold(const pf_bslot *__restrict__ T)
{
    bs_fold<WHICH>(T, blockIdx.x);
}

// ------------------------------------------ persistent batched iterations
// Grid-wide barrier of a cooperative launch (every workgroup resident): an
// arrival counter and a generation word.  The last workgroup to arrive
// resets the counter and advances the generation; the others wait for the
// generation to move, at most about 2^26 polls (seconds).  A wait that runs
// out sets the error word, which every other waiting workgroup sees, and the
// barrier returns false: the kernel then ends in every workgroup.  Agent-
// scope release/acquire make each workgroup's writes visible to the others
// (the L2 of every XCD) across the barrier.
struct pf_gbar {
    unsigned count;
    unsigned gen;
    int error;
    int iters;   // iterations the persistent kernel completed
};

__device__ bool grid_barrier(pf_gbar *g, unsigned nb)
{
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int good = 1;
        const unsigned gen = __hip_atomic_load(&g->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned arrived =
            __hip_atomic_fetch_add(&g->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (arrived == nb) {
            __hip_atomic_store(&g->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&g->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long polls = 0;
            while (__hip_atomic_load(&g->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (__hip_atomic_load(&g->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++polls > (1L << 26)) {
                    __hip_atomic_store(&g->error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

// The iteration of every solve, repeated until every solve has stopped or
// max_iters iterations have run, in one cooperative launch: each phase's
// blocks (the same blocks, with the same work, as the launches per kernel)
// are dealt to the resident workgroups in rounds, XCD-contiguous within a
// round, with a grid barrier between phases.  The true-residual phases run
// only in iterations where some solve asks for a replacement.
template <int NT, int KIND, typename OUT, bool CHECON, bool REF>
__global__ __launch_bounds__(PF_BLOCK) void kb_persistent(const pf_bslot *__restrict__ T, const long *__restrict__ est,
                                                         const int *__restrict__ qst, int ns, long total_eb,
                                                         long total_ebr, int total_qb, int max_iters, pf_gbar *bar)
{
    const unsigned nb = gridDim.x;
    const long lbx = xcd_block(blockIdx.x, nb);
    int it = 0;
    for (; it < max_iters; ++it) {
        bool any = false;
        for (int s = 0; s < ns; ++s) any |= !as_global(T[s].st)->frozen;   // the same in every workgroup
        if (!any) break;
        for (long r0 = 0; r0 < total_eb; r0 += nb) {
            const long gb = r0 + lbx;
            if (gb < total_eb) bs_ebe<NT, KIND, OUT, false>(T, est, ns, gb, RUN_LIVE);
            __builtin_amdgcn_wave_barrier();
        }
        if (!grid_barrier(bar, nb)) return;
        for (int b = blockIdx.x; b < total_qb; b += nb) bs_scatter_pu<OUT>(T, qst, ns, b);
        if (!grid_barrier(bar, nb)) return;
        for (int s = blockIdx.x; s < ns; s += nb) bs_fold<0>(T, s);
        if (!grid_barrier(bar, nb)) return;
        for (int b = blockIdx.x; b < total_qb; b += nb) bs_update<CHECON>(T, qst, ns, b);
        if (!grid_barrier(bar, nb)) return;
        for (int s = blockIdx.x; s < ns; s += nb) bs_fold<CHECON ? 1 : 2>(T, s);
        if (!grid_barrier(bar, nb)) return;
        if constexpr (REF) {
            bool due = false;
            for (int s = 0; s < ns; ++s) due |= !skip(as_global(T[s].st), RUN_REFINE);
            if (due) {
                bs_ebe_refine<NT>(T, est + PF_BATCH_MAX, ns, total_ebr, blockIdx.x, nb);
                if (!grid_barrier(bar, nb)) return;
                for (int b = blockIdx.x; b < total_qb; b += nb) bs_scatter_res(T, qst, ns, b);
                if (!grid_barrier(bar, nb)) return;
                for (int s = blockIdx.x; s < ns; s += nb) bs_fold<3>(T, s);
                if (!grid_barrier(bar, nb)) return;
            }
        }
        for (int b = blockIdx.x; b < total_qb; b += nb) bs_pupdate(T, qst, ns, b);
        if (!grid_barrier(bar, nb)) return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) bar->iters = it;
}

// Copies of every solve's state (pf_batch_query's snapshot).
__global__ void kb_snapshot(const pf_bslot *__restrict__ T, int ns, pf_state *__restrict__ out)
{
    const int s = blockIdx.x;
    if (s < ns && threadIdx.x == 0) out[s] = *as_global(T[s].st);
}

}  // namespace

// ------------------------------------------------ allocation-free K x
// pf_apply's buffers (padded x, y, element products) in a file-scope device
// array instead of allocations.  In host code `scratch` names a host-side
// stand-in, not device memory: kernels and copies take the per-device
// address that hipGetSymbolAddress returns (g_scratch_dev).
static __device__ __align__(256) unsigned char scratch[PF_NOALLOC_SCRATCH_BYTES];
static constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

int pf_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return PF_ERR_NODEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return PF_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return pf_set_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)",
                                __FILE__, __LINE__);
    }
    g_scratch_dev[device] = ptr;
    return PF_OK;
}

extern "C" {

int pf_dev_copy(pf_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (bytes && (!dst || !src)) || bytes % 16) return PF_ERR_ARG;
    if (!bytes) return PF_OK;
    const long n16 = (long)(bytes / 16);
    const int grid = (int)std::min<long>((n16 + COPY_CHUNK - 1) / COPY_CHUNK, 1024);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, (const u4_t *)src, (u4_t *)dst, n16);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

static int vec_op(pf_ctx *c, int op, double *w, const double *x, const double *y, double a, int64_t n)
{
    if (!c || n < 0 || (n && !w)) return PF_ERR_ARG;
    if (n && op != 0 && !x) return PF_ERR_ARG;
    if (n && (op == 1 || op == 2) && !y) return PF_ERR_ARG;
    if (!n) return PF_OK;
    int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_vec, dim3(grid), dim3(256), 0, c->stream, op, w, x, y, a, (long)n);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_vec_fill(pf_ctx *c, double *x, double value, int64_t n) { return vec_op(c, 0, x, nullptr, nullptr, value, n); }
int pf_vec_waxpy(pf_ctx *c, double *w, const double *x, const double *y, double a, int64_t n)
{
    return vec_op(c, 1, w, x, y, a, n);
}
int pf_vec_mul(pf_ctx *c, double *z, const double *x, const double *y, int64_t n) { return vec_op(c, 2, z, x, y, 1.0, n); }
int pf_vec_recip(pf_ctx *c, double *y, const double *x, int64_t n) { return vec_op(c, 3, y, x, nullptr, 1.0, n); }

int pf_gather(pf_ctx *c, const double *p, double *pmul)
{
    if (!c || !p || !pmul) return PF_ERR_ARG;
    if (!c->g_g) return PF_ERR_STATE;
    const long n = (long)c->ntot * c->nels;
    hipLaunchKernelGGL(k_gather, dim3(4096), dim3(256), 0, c->stream, c->g_g, p, pmul, n);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_elem_matvec(pf_ctx *c, const double *pmul, double *utemp)
{
    if (!c || !pmul || !utemp) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    return launch_ebe_exact<false>(c, pmul, utemp);
}

int pf_scatter(pf_ctx *c, double *u, const double *utemp)
{
    if (!c || !u || !utemp) return PF_ERR_ARG;
    if (!c->csr_ptr) return PF_ERR_STATE;
    hipLaunchKernelGGL(k_scatter_prim, dim3(pf_grid_for(c->neq)), dim3(256), 0, c->stream, c->csr_ptr, c->csr_slot,
                       utemp, u, c->neq, 1);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_scatter_noadd(pf_ctx *c, const double *utemp, double *u)
{
    if (!c || !u || !utemp) return PF_ERR_ARG;
    if (!c->csr_ptr) return PF_ERR_STATE;
    hipLaunchKernelGGL(k_scatter_prim, dim3(pf_grid_for(c->neq)), dim3(256), 0, c->stream, c->csr_ptr, c->csr_slot,
                       utemp, u, c->neq, 0);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_dot_product_p(pf_ctx *c, const double *a, const double *b, int64_t n, double *out)
{
    if (!c || !out || n < 0 || (n && (!a || !b))) return PF_ERR_ARG;
    double o[2];
    int rc = reduce_prim(c, a, const_cast<double *>(b), n, 0, o);
    if (!rc) *out = o[0];
    return rc;
}

int pf_sum_p(pf_ctx *c, const double *a, int64_t n, double *out)
{
    if (!c || !out || n < 0 || (n && !a)) return PF_ERR_ARG;
    double o[2];
    int rc = reduce_prim(c, a, nullptr, n, 1, o);
    if (!rc) *out = o[0];
    return rc;
}

int pf_dot_order_set(pf_ctx *c, int order)
{
    if (!c || (order != PF_DOT_TREE && order != PF_DOT_SEQ)) return PF_ERR_ARG;
    c->dot_order = order;
    return PF_OK;
}

int pf_dot_order_get(pf_ctx *c, int *order)
{
    if (!c || !order) return PF_ERR_ARG;
    *order = c->dot_order;
    return PF_OK;
}

int pf_scatter_mode_set(pf_ctx *c, int mode)
{
    if (!c || (mode != PF_SCATTER_SLOTS && mode != PF_SCATTER_RANGES)) return PF_ERR_ARG;
    if (mode != c->scatter_mode) {
        pf_graph_drop(c);
        c->pcg_ready = false;   // pf_pcg_begin again (it builds the ranges)
    }
    c->scatter_mode = mode;
    return PF_OK;
}

int pf_scatter_mode_get(pf_ctx *c, int *mode)
{
    if (!c || !mode) return PF_ERR_ARG;
    *mode = c->scatter_mode;
    return PF_OK;
}

int pf_checon_par(pf_ctx *c, const double *xnew, double tol, int *converged, double *x, int64_t n)
{
    if (!c || !converged || n < 0 || (n && (!xnew || !x))) return PF_ERR_ARG;
    double o[2];
    int rc = reduce_prim(c, xnew, x, n, 2, o);
    if (!rc) *converged = (o[1] / o[0]) <= tol;
    return rc;
}

// pf_pcg_begin, with x starting at x0 (device, neq) when x0 is given.
static int pcg_begin(pf_ctx *c, const double *r0, int r0_on_device, const double *x0, const pf_pcg_opts *opts)
{
    if (!c || !r0 || !opts) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    if (opts->limit < 1 || !(opts->tol >= 0.0)) return PF_ERR_ARG;
    if (opts->stop_mode != PF_STOP_CHECON && opts->stop_mode != PF_STOP_RELRES) return PF_ERR_ARG;
    if (opts->precision == PF_PREC_FP64) {
        if (!c->K64) return PF_ERR_ARG;   // FP64 iterations need FP64 storage
    } else if (opts->precision == PF_PREC_SPLIT) {
        if (!c->K32hi) return PF_ERR_ARG;
    } else if (opts->precision == PF_PREC_MIXED || opts->precision == PF_PREC_MIXED64) {
        if (!c->K32hi || opts->stop_mode != PF_STOP_RELRES) return PF_ERR_ARG;
        if (opts->refine_delta < 0.0 || opts->refine_delta >= 1.0 || opts->refine_every < 0) return PF_ERR_ARG;
    } else {
        return PF_ERR_ARG;
    }
    PF_HIP(hipSetDevice(c->device));
    // fl32(p) beside p where the gathered vector is a large share of an
    // element's bytes (hex8 Laplace: 64 of 256; hex20: 480 of 8160)
    if (opts->precision == PF_PREC_MIXED && c->ntot <= 24 && !c->p32) {
        pf_graph_drop(c);
        const int rc = pf_alloc(c, (void **)&c->p32, ((size_t)c->neq + 1) * sizeof(float));
        if (rc) return rc;
    }
    const bool ranges = opts->precision == PF_PREC_MIXED && c->scatter_mode == PF_SCATTER_RANGES;
    if (ranges && c->dot_order != PF_DOT_TREE) return PF_ERR_ARG;   // p.u is summed per range
    if (ranges && !c->rng) {
        pf_graph_drop(c);
        const int rc = pf_ranges_build(c);
        if (rc) return rc;
    }
    c->opts = *opts;
    PF_HIP(hipMemcpyAsync(c->r + 1, r0, sizeof(double) * c->neq,
                          r0_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    pf_state s{};
    s.tol = opts->tol;
    s.limit = opts->limit;
    s.stop_mode = opts->stop_mode;
    s.precision = opts->precision;
    const double delta = opts->refine_delta > 0.0 ? opts->refine_delta : 1e-2;
    s.delta2 = delta * delta;
    s.refine_every = opts->refine_every;
    s.pfresh = 1;
    *c->state_host = s;
    PF_HIP(hipMemcpyAsync(c->state, c->state_host, sizeof(pf_state), hipMemcpyHostToDevice, c->stream));
    const int G = pf_grid_for(c->neq);
    float *p32 = opts->precision == PF_PREC_MIXED && c->ntot <= 24 ? c->p32 : nullptr;
    hipLaunchKernelGGL(k_pcg_init, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->x, c->r, c->d, c->p, p32, c->dinv, c->diag,
                       c->f, x0, c->neq, c->precon_given ? 1 : 0, c->partials);
    const int Gi = seq_reduce(c, G, c->r + 1, c->d + 1, PF_Q_RD, c->r + 1, c->r + 1, PF_Q_RR, false, c->neq,
                              nullptr, RUN_ALWAYS);
    hipLaunchKernelGGL(k_fold_init, dim3(1), dim3(PF_BLOCK), 0, c->stream, c->partials, Gi, c->state);
    PF_HIP(hipGetLastError());
    // The host mirror is read back by pf_pcg_poll; wait so it is not reused
    // while the copy above is still in flight.
    PF_HIP(hipStreamSynchronize(c->stream));
    c->pcg_ready = true;
    c->query_pending = false;
    return PF_OK;
}

int pf_pcg_begin(pf_ctx *c, const double *r0, int r0_on_device, const pf_pcg_opts *opts)
{
    return pcg_begin(c, r0, r0_on_device, nullptr, opts);
}

int pf_pcg_run(pf_ctx *c, int n)
{
    if (!c || n < 0) return PF_ERR_ARG;
    if (!c->pcg_ready) return PF_ERR_STATE;
    const int G = pf_grid_for(c->neq);
    const bool mixed = c->opts.precision == PF_PREC_MIXED;
    const bool wide = c->opts.precision == PF_PREC_MIXED64;
    const bool split = c->opts.precision == PF_PREC_SPLIT;
    float *p32 = mixed && c->ntot <= 24 ? c->p32 : nullptr;
    // hex8 with FP64 products: row-major (row i of every element side by
    // side, read back through the matching slot list; the scatter's reads of
    // consecutive equations then fall on consecutive words).  The FP32
    // products of the mixed path stay in element order: their 4-byte
    // row-major writes cost the product more than the scatter saves.
    const int rm = c->csr_slot_t && !mixed ? 1 : 0;
    const int *slot = rm ? c->csr_slot_t : c->csr_slot;
    const bool ranges = mixed && c->scatter_mode == PF_SCATTER_RANGES;
    if (ranges && !c->rng) return PF_ERR_STATE;   // pf_pcg_begin builds them
    hipEvent_t t0, t1;
    for (int it = 0; it < n; ++it) {
        int rc;
        bool tm = pf_timing_begin(c, PF_K_EBE, &t0, &t1);
        if (tm) PF_HIP(hipEventRecord(t0, c->stream));
        if (ranges)
            rc = pf_ranges_iteration(c);   // p, the products, their sums into u, p.u per range
        else if (mixed)
            rc = launch_ebe<KIND_F32, true, float>(c, c->p, c->utemp32, c->state, RUN_LIVE, nullptr, p32, rm);
        else if (wide)
            rc = launch_ebe<KIND_WIDE, true, double>(c, c->p, c->utemp, c->state, RUN_LIVE, nullptr, nullptr, rm);
        else if (split)
            rc = launch_ebe<KIND_ACC, true, double>(c, c->p, c->utemp, c->state, RUN_LIVE, nullptr, nullptr, rm);
        else
            rc = launch_ebe<KIND_F64, true, double>(c, c->p, c->utemp, c->state, RUN_LIVE, nullptr, nullptr, rm);
        if (rc) return rc;
        if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        tm = pf_timing_begin(c, PF_K_SCATTER, &t0, &t1);
        if (tm) PF_HIP(hipEventRecord(t0, c->stream));
        if (ranges) {
            rc = pf_ranges_fold(c);
            if (rc) return rc;
        } else if (mixed)
            hipLaunchKernelGGL(k_scatter_pu<float>, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->csr_ptr, slot,
                               c->utemp32, c->fixstore, c->p, c->u, c->neq, c->fix_zero ? 1 : 0, c->partials,
                               c->state);
        else
            hipLaunchKernelGGL(k_scatter_pu<double>, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->csr_ptr,
                               slot, c->utemp, c->fixstore, c->p, c->u, c->neq, c->fix_zero ? 1 : 0,
                               c->partials, c->state);
        if (!ranges) {
            const int Ga = seq_reduce(c, G, c->p + 1, c->u + 1, PF_Q_PU, nullptr, nullptr, 0, false, c->neq,
                                      c->state, RUN_LIVE);
            hipLaunchKernelGGL(k_fold_alpha, dim3(1), dim3(PF_BLOCK), 0, c->stream, c->partials, Ga, c->state);
        }
        if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        tm = pf_timing_begin(c, PF_K_UPDATE, &t0, &t1);
        if (tm) PF_HIP(hipEventRecord(t0, c->stream));
        if (c->opts.stop_mode == PF_STOP_CHECON) {
            hipLaunchKernelGGL(k_update<true>, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->x, c->r, c->d, c->p, c->u,
                               c->dinv, c->neq, c->partials, c->state);
            const int Gu = seq_reduce(c, G, c->r + 1, c->d + 1, PF_Q_RD, nullptr, nullptr, 0, true, c->neq,
                                      c->state, RUN_LIVE);
            hipLaunchKernelGGL(k_fold_update<true>, dim3(1), dim3(PF_BLOCK), 0, c->stream, c->partials, Gu,
                               c->state);
        } else {
            hipLaunchKernelGGL(k_update<false>, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->x, c->r, c->d, c->p, c->u,
                               c->dinv, c->neq, c->partials, c->state);
            const int Gu = seq_reduce(c, G, c->r + 1, c->d + 1, PF_Q_RD, c->r + 1, c->r + 1, PF_Q_RR, false,
                                      c->neq, c->state, RUN_LIVE);
            hipLaunchKernelGGL(k_fold_update<false>, dim3(1), dim3(PF_BLOCK), 0, c->stream, c->partials, Gu,
                               c->state);
        }
        if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        if (mixed || wide) {
            // True-residual replacement; the kernels do nothing unless the
            // fold above flagged this iteration.
            tm = pf_timing_begin(c, PF_K_REFINE, &t0, &t1);
            if (tm) PF_HIP(hipEventRecord(t0, c->stream));
            rc = launch_ebe<KIND_ACC, true, double>(c, c->x, c->utemp, c->state, RUN_REFINE, nullptr, nullptr, rm);
            if (rc) return rc;
            hipLaunchKernelGGL(k_scatter_res, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->csr_ptr, slot,
                               c->utemp, c->fixstore, c->x, c->f, c->r, c->d, c->dinv, c->neq, c->partials,
                               c->state);
            const int Gr = seq_reduce(c, G, c->r + 1, c->d + 1, PF_Q_RD, c->r + 1, c->r + 1, PF_Q_RR, false,
                                      c->neq, c->state, RUN_REFINE);
            hipLaunchKernelGGL(k_fold_refine, dim3(1), dim3(PF_BLOCK), 0, c->stream, c->partials, Gr, c->state);
            if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        }
        if (!ranges || !pf_ranges_form_p(c)) {   // else the next range kernels form p
            tm = pf_timing_begin(c, PF_K_PUPDATE, &t0, &t1);
            if (tm) PF_HIP(hipEventRecord(t0, c->stream));
            hipLaunchKernelGGL(k_pupdate, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->p, p32, c->d, c->neq, c->state);
            if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        }
        PF_HIP(hipGetLastError());
    }
    return PF_OK;
}

int pf_pcg_run_graph(pf_ctx *c, int n)
{
    if (!c || n < 1) return PF_ERR_ARG;
    if (!c->pcg_ready) return PF_ERR_STATE;
    if (c->timing) return pf_pcg_run(c, n);   // timing records an event per launch
    if (!c->graph_exec || c->graph_iters != n || c->graph_precision != c->opts.precision ||
        c->graph_stop != c->opts.stop_mode || c->graph_dot != c->dot_order || c->graph_scatter != c->scatter_mode) {
        pf_graph_drop(c);
        PF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = pf_pcg_run(c, n);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : pf_set_hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
        }
        const hipError_t ei = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            c->graph_exec = nullptr;
            return pf_set_hip_error(ei, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        c->graph = g;
        c->graph_iters = n;
        c->graph_precision = c->opts.precision;
        c->graph_stop = c->opts.stop_mode;
        c->graph_dot = c->dot_order;
        c->graph_scatter = c->scatter_mode;
    }
    PF_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    return PF_OK;
}

static void fill_result(const pf_state &s, pf_pcg_result *res)
{
    res->iters = s.iters;
    res->converged = s.converged;
    res->crit = s.crit;
    res->rr0 = s.rr0;
    res->rr = s.rr;
    res->refinements = s.refinements;
    res->status = s.breakdown ? PF_ERR_BREAKDOWN : (s.converged ? PF_OK : PF_NOT_CONVERGED);
}

int pf_pcg_query(pf_ctx *c, pf_pcg_result *res, int *ready)
{
    if (!c || !res || !ready) return PF_ERR_ARG;
    if (!c->pcg_ready) return PF_ERR_STATE;
    *ready = 0;
    if (!c->query_pending) {
        PF_HIP(hipMemcpyAsync(c->state_query, c->state, sizeof(pf_state), hipMemcpyDeviceToHost, c->stream));
        PF_HIP(hipEventRecord(c->ev_query, c->stream));
        c->query_pending = true;
    }
    const hipError_t q = hipEventQuery(c->ev_query);
    if (q == hipErrorNotReady) return PF_OK;
    if (q != hipSuccess) return pf_set_hip_error(q, "hipEventQuery", __FILE__, __LINE__);
    c->query_pending = false;
    fill_result(*c->state_query, res);
    *ready = 1;
    return PF_OK;
}

int pf_pcg_timing_enable(pf_ctx *c, int on)
{
    if (!c) return PF_ERR_ARG;
    PF_HIP(hipStreamSynchronize(c->stream));
    c->timing = on != 0;
    c->ev_next = 0;
    for (auto &v : c->ev_used) v.clear();
    return PF_OK;
}

int pf_pcg_timing_get(pf_ctx *c, double ms[PF_TIMING_SLOTS], int counts[PF_TIMING_SLOTS])
{
    if (!c || !ms || !counts) return PF_ERR_ARG;
    PF_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < PF_NKERNELS; ++k) {
        double s = 0.0;
        for (auto &pr : c->ev_used[k]) {
            float t = 0.f;
            PF_HIP(hipEventElapsedTime(&t, c->ev_pool[pr.first], c->ev_pool[pr.second]));
            s += t;
        }
        ms[k] = s;
        counts[k] = (int)c->ev_used[k].size();
    }
    return PF_OK;
}

int pf_pcg_poll(pf_ctx *c, pf_pcg_result *res)
{
    if (!c || !res) return PF_ERR_ARG;
    if (!c->pcg_ready) return PF_ERR_STATE;
    PF_HIP(hipMemcpyAsync(c->state_host, c->state, sizeof(pf_state), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    fill_result(*c->state_host, res);
    return PF_OK;
}

int pf_pcg_x(pf_ctx *c, double *x, int on_device)
{
    if (!c || !x) return PF_ERR_ARG;
    if (!c->pcg_ready) return PF_ERR_STATE;
    PF_HIP(hipMemcpyAsync(x, c->x + 1, sizeof(double) * c->neq,
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_pcg_r(pf_ctx *c, double *r, int on_device)
{
    if (!c || !r) return PF_ERR_ARG;
    if (!c->pcg_ready) return PF_ERR_STATE;
    PF_HIP(hipMemcpyAsync(r, c->r + 1, sizeof(double) * c->neq,
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

// Iterations in chunks of check_every (default 16) until the state freezes.
static int pcg_finish(pf_ctx *c, const pf_pcg_opts *opts, pf_pcg_result *R)
{
    const int chunk = opts->check_every > 0 ? opts->check_every : 16;
    int rc = pf_pcg_poll(c, R);
    while (!rc && !c->state_host->frozen) {
        int n = chunk;
        const int left = opts->limit - R->iters;
        if (n > left) n = left;
        if (n < 1) n = 1;
        rc = pf_pcg_run(c, n);
        if (!rc) rc = pf_pcg_poll(c, R);
    }
    return rc;
}

int pf_pcg_solve(pf_ctx *c, const double *r0, double *x, const pf_pcg_opts *opts, pf_pcg_result *res)
{
    if (!c || !r0 || !x || !opts) return PF_ERR_ARG;
    int rc = pf_pcg_begin(c, r0, 0, opts);
    if (rc) return rc;
    pf_pcg_result tmp{};
    pf_pcg_result *R = res ? res : &tmp;
    rc = pcg_finish(c, opts, R);
    if (rc) return rc;
    rc = pf_pcg_x(c, x, 0);
    if (rc) return rc;
    return R->status;
}

int pf_pcg_solve_warm(pf_ctx *c, const double *r0, const double *x0, double *xnew, int fixed_mode,
                      const pf_pcg_opts *opts, pf_pcg_result *res)
{
    if (!c || !r0 || !x0 || !xnew || !opts || (fixed_mode != 0 && fixed_mode != 1)) return PF_ERR_ARG;
    // no residual replacement: its true residual would apply the fixed rows
    // that the caller's r0 = loads - K x0 leaves out
    if (opts->precision != PF_PREC_FP64 && opts->precision != PF_PREC_SPLIT) return PF_ERR_ARG;
    int rc = pcg_begin(c, r0, 1, x0, opts);
    if (rc) return rc;
    pf_pcg_result tmp{};
    pf_pcg_result *R = res ? res : &tmp;
    c->fix_zero = fixed_mode == 0;   // read by the launches of pf_pcg_run below
    rc = pcg_finish(c, opts, R);
    c->fix_zero = false;
    if (rc) return rc;
    rc = pf_pcg_x(c, xnew, 1);
    if (rc) return rc;
    return R->status;
}

// y = K x through gather / element products / ascending scatter in FP64 (the
// stored matrices, or K64b), with or without the p123 fixed-freedom rows.
static int apply_host(pf_ctx *c, const double *x_host, double *y_host, const double *Kalt, bool fixed_rows)
{
    PF_HIP(hipSetDevice(c->device));
    const size_t nv = (size_t)c->neq + 1, ns = (size_t)c->ntot * c->nels;
    double *xe = nullptr, *y = nullptr, *ut = nullptr;
    int rc = pf_alloc(c, (void **)&xe, nv * sizeof(double));
    if (!rc) rc = pf_alloc(c, (void **)&y, nv * sizeof(double));
    if (!rc) rc = pf_alloc(c, (void **)&ut, ns * sizeof(double));
    if (!rc) {
        const double zero = 0.0;
        hipError_t e = hipMemcpyAsync(xe, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(xe + 1, x_host, c->neq * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = pf_set_hip_error(e, "pf_apply upload", __FILE__, __LINE__);
    }
    if (!rc) {
        if (Kalt) rc = launch_ebe<KIND_F64, true, double>(c, xe, ut, nullptr, RUN_ALWAYS, Kalt);
        else rc = launch_ebe_exact<true>(c, xe, ut);
    }
    if (!rc) {
        hipLaunchKernelGGL(k_scatter_fix, dim3(pf_grid_for(c->neq)), dim3(256), 0, c->stream, c->csr_ptr,
                           c->csr_slot, ut, fixed_rows ? c->fixstore : (const double *)nullptr, xe, y, c->neq);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(y_host, y, c->neq * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = pf_set_hip_error(e, "pf_apply", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(c->stream);
    pf_free(c, xe, nv * sizeof(double));
    pf_free(c, y, nv * sizeof(double));
    pf_free(c, ut, ns * sizeof(double));
    return rc;
}

int pf_apply(pf_ctx *c, const double *x_host, double *y_host)
{
    if (!c || !x_host || !y_host) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    return apply_host(c, x_host, y_host, nullptr, true);
}

int pf_apply_noalloc(pf_ctx *c, const double *x_host, double *y_host)
{
    if (!c || !x_host || !y_host) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    auto up256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t nv = (size_t)c->neq + 1, ns = (size_t)c->ntot * c->nels;
    const size_t bv = up256(nv * sizeof(double)), bs = up256(ns * sizeof(double));
    if (2 * bv + bs > PF_NOALLOC_SCRATCH_BYTES) return PF_ERR_TOO_LARGE;
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c->device >= 0 && c->device < kMaxDevices) base = (unsigned char *)g_scratch_dev[c->device];
    }
    if (!base) return PF_ERR_STATE;   // pf_create looks the address up; not reached after a good create
    double *xe = (double *)base, *y = (double *)(base + bv), *ut = (double *)(base + 2 * bv);
    PF_HIP(hipSetDevice(c->device));
    // xe[0] must read 0 (a restrained freedom gathers slot 0); the scratch
    // keeps the previous call's contents, so it is written on every call
    static const double zero = 0.0;
    PF_HIP(hipMemcpyAsync(xe, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream));
    PF_HIP(hipMemcpyAsync(xe + 1, x_host, c->neq * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = launch_ebe_exact<true>(c, xe, ut);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scatter_fix, dim3(pf_grid_for(c->neq)), dim3(256), 0, c->stream, c->csr_ptr, c->csr_slot,
                       ut, (const double *)c->fixstore, xe, y, c->neq);
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(y_host, y, c->neq * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_apply_raw(pf_ctx *c, int which, const double *x_host, double *y_host)
{
    if (!c || !x_host || !y_host || (which != 0 && which != 1)) return PF_ERR_ARG;
    if (!c->have_K || (which == 1 && !c->K64b)) return PF_ERR_STATE;
    return apply_host(c, x_host, y_host, which == 1 ? c->K64b : nullptr, false);
}

}  // extern "C"

// ------------------------------------------------- batched iterations
// pf_batch: one stream, a device table of up to PF_BATCH_MAX solves, and
// the captured iteration chunk for the current table.
struct pf_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_done = nullptr;                  // end of the last enqueued run
    hipEvent_t ev_table = nullptr;                 // the last table upload has left the host buffer
    hipEvent_t ev_snap = nullptr;
    std::vector<hipEvent_t> ev_ctx;                // one per solve: its context's stream reached the run
    pf_bslot *T = nullptr;                         // device table
    long *est = nullptr;                           // device: first element block per solve (iteration
                                                   // kind), then the same for the FP64 hi+lo products
    int *qst = nullptr;                            // device: first equation block per solve
    pf_bslot *T_host = nullptr;                    // pinned staging of the three (one allocation)
    long *est_host = nullptr;
    int *qst_host = nullptr;
    std::vector<pf_bslot> cur;                     // the table on the device
    std::vector<long> cur_est;                     // iteration-kind starts, then ACC-kind starts
    std::vector<int> cur_qst;
    int ns = 0, ntot = 0, precision = -1, stop = -1;
    long total_eb = 0, total_ebr = 0;              // element blocks: iteration kind, true-residual kind
    int total_qb = 0;
    pf_state *snap = nullptr, *snap_host = nullptr;
    int snap_n = 0;
    bool snap_pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_iters = 0;
    // persistent runs: the grid barrier (device) and its copy taken with
    // each snapshot, the device's CU count and cooperative-launch support
    pf_gbar *bar = nullptr, *bar_host = nullptr;
    int cus = 0, coop = 0;
    bool persistent_used = false;
    // solves of at least solo_elems elements (0: none) run their own
    // launches, in the batch's stream and graph, instead of a share of the
    // batched ones: a large solve fills the device by itself, and its own
    // kernels skip the slot lookup the batched ones do per workgroup
    long solo_elems = 0;
    std::vector<pf_ctx *> solo;
    bool stale = false;   // the next run rebuilds the table (and drops the graph)
};

static const size_t kBatchTable = sizeof(pf_bslot) * PF_BATCH_MAX + 2 * sizeof(long) * PF_BATCH_MAX +
                                  sizeof(int) * PF_BATCH_MAX;

static void batch_graph_drop(pf_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->graph_iters = 0;
}

// Builds the table of the given contexts, uploads it when it changed and
// orders the batch stream after each context's stream.
static int batch_prepare(pf_batch *b, int nctx, pf_ctx *const *ctxs, bool allow_solo = true)
{
    if (!b || nctx < 1 || nctx > PF_BATCH_MAX || !ctxs) return PF_ERR_ARG;
    const pf_ctx *c0 = ctxs[0];
    if (!c0) return PF_ERR_ARG;
    std::vector<pf_bslot> T(nctx);
    std::vector<long> est(2 * (size_t)nctx);
    std::vector<int> qst(nctx);
    long eb = 0;
    int qb = 0;
    for (int k = 0; k < nctx; ++k) {
        pf_ctx *c = ctxs[k];
        if (!c || c->device != b->device) return PF_ERR_ARG;
        for (int j = 0; j < k; ++j)
            if (ctxs[j] == c) return PF_ERR_ARG;   // a solve appears once
        if (!c->pcg_ready || c->timing) return PF_ERR_STATE;
        if (c->dot_order != PF_DOT_TREE || c->ntot != c0->ntot || c->opts.precision != c0->opts.precision ||
            c->opts.stop_mode != c0->opts.stop_mode)
            return PF_ERR_ARG;
        if (c->opts.precision == PF_PREC_MIXED && c->scatter_mode != PF_SCATTER_SLOTS) return PF_ERR_ARG;
        const bool mixed = c->opts.precision == PF_PREC_MIXED;
        const int rm = c->csr_slot_t && !mixed ? 1 : 0;
        pf_bslot &S = T[k];
        S.K64 = c->K64;
        S.Khi = c->K32hi;
        S.Klo = c->K32lo;
        S.g_g = c->g_g;
        S.csr_ptr = c->csr_ptr;
        S.slot = rm ? c->csr_slot_t : c->csr_slot;
        S.x = c->x; S.r = c->r; S.p = c->p; S.d = c->d; S.u = c->u; S.dinv = c->dinv; S.f = c->f;
        S.utemp = c->utemp;
        S.p32 = mixed && c->ntot <= 24 ? c->p32 : nullptr;
        S.utemp32 = c->utemp32;
        S.fixstore = c->fixstore;
        S.partials = c->partials;
        S.st = c->state;
        S.kstride = pf_k64_estride(c);
        S.nels = c->nels;
        S.neq = c->neq;
        S.G = pf_grid_for(c->neq);
        S.fix_zero = c->fix_zero ? 1 : 0;
        S.rowmajor = rm;
        S.solo = allow_solo && b->solo_elems > 0 && c->nels >= b->solo_elems ? 1 : 0;
        qst[k] = qb;
        if (!S.solo) qb += S.G;
        if (mixed && !c->utemp32) return PF_ERR_STATE;
    }
    // elements per logical block of the iteration's product kind and of the
    // FP64 hi+lo products (the true residual)
    int EPB = 0, EPBR = 0;
    const int prec = c0->opts.precision;
    auto epb = [&](auto nt) {
        constexpr int NT = decltype(nt)::value;
        EPB = prec == PF_PREC_MIXED ? Lanes<NT, KIND_F32>::EPB
            : prec == PF_PREC_FP64 ? Lanes<NT, KIND_F64>::EPB : Lanes<NT, KIND_ACC>::EPB;
        EPBR = Lanes<NT, KIND_ACC>::EPB;
        static_assert(Lanes<NT, KIND_ACC>::EPB == Lanes<NT, KIND_WIDE>::EPB, "one lane layout");
    };
    switch (c0->ntot) {
    case 60: epb(std::integral_constant<int, 60>()); break;
    case 24: epb(std::integral_constant<int, 24>()); break;
    case 8: epb(std::integral_constant<int, 8>()); break;
    default: return PF_ERR_ARG;
    }
    long ebr = 0;
    std::vector<pf_ctx *> solo;
    for (int k = 0; k < nctx; ++k) {
        est[k] = eb;
        est[nctx + k] = ebr;
        if (T[k].solo) {
            solo.push_back(ctxs[k]);
            continue;
        }
        eb += (T[k].nels + EPB - 1) / EPB;
        ebr += (T[k].nels + EPBR - 1) / EPBR;
    }
    PF_HIP(hipSetDevice(b->device));
    const bool same = !b->stale && nctx == b->ns &&
                      std::memcmp(T.data(), b->cur.data(), sizeof(pf_bslot) * nctx) == 0 &&
                      est == b->cur_est && qst == b->cur_qst;
    if (!same) {
        PF_HIP(hipEventSynchronize(b->ev_table));   // the staging buffer is free again
        std::memcpy(b->T_host, T.data(), sizeof(pf_bslot) * nctx);
        std::memcpy(b->est_host, est.data(), sizeof(long) * nctx);
        std::memcpy(b->est_host + PF_BATCH_MAX, est.data() + nctx, sizeof(long) * nctx);
        std::memcpy(b->qst_host, qst.data(), sizeof(int) * nctx);
        PF_HIP(hipMemcpyAsync(b->T, b->T_host, kBatchTable, hipMemcpyHostToDevice, b->stream));
        PF_HIP(hipEventRecord(b->ev_table, b->stream));
        // a captured graph reads the batched solves from the table, but
        // holds the solo solves' launches with their own buffers
        if (eb != b->total_eb || ebr != b->total_ebr || qb != b->total_qb || nctx != b->ns || c0->ntot != b->ntot ||
            c0->opts.precision != b->precision || c0->opts.stop_mode != b->stop || !solo.empty() ||
            !b->solo.empty() || b->stale)
            batch_graph_drop(b);
        b->solo = solo;
        b->stale = false;
        b->cur = T;
        b->cur_est = est;
        b->cur_qst = qst;
        b->ns = nctx;
        b->total_eb = eb;
        b->total_ebr = ebr;
        b->total_qb = qb;
        b->ntot = c0->ntot;
        b->precision = c0->opts.precision;
        b->stop = c0->opts.stop_mode;
    }
    // the solves' own streams first (their set-up, pf_pcg_begin)
    for (int k = 0; k < nctx; ++k) {
        PF_HIP(hipEventRecord(b->ev_ctx[k], ctxs[k]->stream));
        PF_HIP(hipStreamWaitEvent(b->stream, b->ev_ctx[k], 0));
    }
    return PF_OK;
}

// After a run: each context's stream waits for it (pf_pcg_x, pf_reset and
// the rest of a context's calls see the batch's work).
static int batch_release(pf_batch *b, int nctx, pf_ctx *const *ctxs)
{
    PF_HIP(hipEventRecord(b->ev_done, b->stream));
    for (int k = 0; k < nctx; ++k) {
        ctxs[k]->query_pending = false;
        PF_HIP(hipStreamWaitEvent(ctxs[k]->stream, b->ev_done, 0));
    }
    return PF_OK;
}

template <int NT, int KIND, typename OUT>
static void batch_launch_ebe(pf_batch *b)
{
    hipLaunchKernelGGL((kb_ebe<NT, KIND, OUT>), dim3((unsigned)b->total_eb), dim3(PF_BLOCK), 0, b->stream, b->T,
                       b->est, b->ns);
}

template <int NT>
static void batch_launch_refine(pf_batch *b)
{
    // usually does nothing: few workgroups that loop over every solve's blocks
    const int grid = (int)std::min<long>(b->total_ebr, 2048);
    hipLaunchKernelGGL(kb_ebe_refine<NT>, dim3(grid), dim3(PF_BLOCK), 0, b->stream, b->T, b->est + PF_BATCH_MAX,
                       b->ns, b->total_ebr);
}

template <int NT>
static void batch_iteration_nt(pf_batch *b)
{
    const int prec = b->precision;
    if (prec == PF_PREC_MIXED) batch_launch_ebe<NT, KIND_F32, float>(b);
    else if (prec == PF_PREC_MIXED64) batch_launch_ebe<NT, KIND_WIDE, double>(b);
    else if (prec == PF_PREC_SPLIT) batch_launch_ebe<NT, KIND_ACC, double>(b);
    else batch_launch_ebe<NT, KIND_F64, double>(b);
    const dim3 gq((unsigned)b->total_qb), gs((unsigned)b->ns), blk(PF_BLOCK);
    if (prec == PF_PREC_MIXED)
        hipLaunchKernelGGL(kb_scatter_pu<float>, gq, blk, 0, b->stream, b->T, b->qst, b->ns);
    else
        hipLaunchKernelGGL(kb_scatter_pu<double>, gq, blk, 0, b->stream, b->T, b->qst, b->ns);
    hipLaunchKernelGGL(kb_fold<0>, gs, blk, 0, b->stream, b->T);
    if (b->stop == PF_STOP_CHECON) {
        hipLaunchKernelGGL(kb_update<true>, gq, blk, 0, b->stream, b->T, b->qst, b->ns);
        hipLaunchKernelGGL(kb_fold<1>, gs, blk, 0, b->stream, b->T);
    } else {
        hipLaunchKernelGGL(kb_update<false>, gq, blk, 0, b->stream, b->T, b->qst, b->ns);
        hipLaunchKernelGGL(kb_fold<2>, gs, blk, 0, b->stream, b->T);
    }
    if (prec == PF_PREC_MIXED || prec == PF_PREC_MIXED64) {
        batch_launch_refine<NT>(b);
        hipLaunchKernelGGL(kb_scatter_res, gq, blk, 0, b->stream, b->T, b->qst, b->ns);
        hipLaunchKernelGGL(kb_fold<3>, gs, blk, 0, b->stream, b->T);
    }
    hipLaunchKernelGGL(kb_pupdate, gq, blk, 0, b->stream, b->T, b->qst, b->ns);
}

template <int NT, int KIND, typename OUT, bool CHECON, bool REF>
static int batch_launch_persistent(pf_batch *b, int max_iters)
{
    const void *kern = reinterpret_cast<const void *>(&kb_persistent<NT, KIND, OUT, CHECON, REF>);
    int per_cu = 0;
    PF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, PF_BLOCK, 0));
    if (per_cu < 1) return PF_ERR_ARG;
    // every workgroup resident (the cooperative launch refuses otherwise),
    // no more than the largest phase has blocks
    const long most = std::max<long>(std::max<long>(b->total_eb, b->total_qb), b->ns);
    int grid = (int)std::min<long>((long)per_cu * b->cus, most);
    if (grid < 1) grid = 1;
    PF_HIP(hipMemsetAsync(b->bar, 0, sizeof(pf_gbar), b->stream));
    const pf_bslot *T = b->T;
    const long *est = b->est;
    const int *qst = b->qst;
    int ns = b->ns, tq = b->total_qb, mi = max_iters;
    long te = b->total_eb, ter = b->total_ebr;
    pf_gbar *bar = b->bar;
    void *args[] = {(void *)&T, (void *)&est, (void *)&qst, (void *)&ns, (void *)&te, (void *)&ter, (void *)&tq,
                    (void *)&mi, (void *)&bar};
    PF_HIP(hipLaunchCooperativeKernel(kern, dim3(grid), dim3(PF_BLOCK), args, 0, b->stream));
    return PF_OK;
}

template <int NT>
static int batch_persistent_nt(pf_batch *b, int max_iters)
{
    const bool ch = b->stop == PF_STOP_CHECON;
    switch (b->precision) {
    case PF_PREC_MIXED:
        return ch ? batch_launch_persistent<NT, KIND_F32, float, true, true>(b, max_iters)
                  : batch_launch_persistent<NT, KIND_F32, float, false, true>(b, max_iters);
    case PF_PREC_MIXED64:
        return ch ? batch_launch_persistent<NT, KIND_WIDE, double, true, true>(b, max_iters)
                  : batch_launch_persistent<NT, KIND_WIDE, double, false, true>(b, max_iters);
    case PF_PREC_SPLIT:
        return ch ? batch_launch_persistent<NT, KIND_ACC, double, true, false>(b, max_iters)
                  : batch_launch_persistent<NT, KIND_ACC, double, false, false>(b, max_iters);
    default:
        return ch ? batch_launch_persistent<NT, KIND_F64, double, true, false>(b, max_iters)
                  : batch_launch_persistent<NT, KIND_F64, double, false, false>(b, max_iters);
    }
}

static int batch_iterations(pf_batch *b, int n)
{
    for (int it = 0; it < n; ++it) {
        // the solo solves: their own iteration (pf_pcg_run's launches) in
        // the batch's stream; no other work of theirs runs meanwhile
        for (pf_ctx *c : b->solo) {
            hipStream_t own = c->stream;
            c->stream = b->stream;
            const int rc = pf_pcg_run(c, 1);
            c->stream = own;
            if (rc) return rc;
        }
        if (b->total_eb > 0) {
            switch (b->ntot) {
            case 60: batch_iteration_nt<60>(b); break;
            case 24: batch_iteration_nt<24>(b); break;
            default: batch_iteration_nt<8>(b); break;
            }
        }
        PF_HIP(hipGetLastError());
    }
    return PF_OK;
}

extern "C" {

int pf_batch_create(int device, pf_batch **out)
{
    if (!out) return PF_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return PF_ERR_NODEVICE;
    }
    if (device < 0 || device >= ndev) return PF_ERR_NODEVICE;
    PF_HIP(hipSetDevice(device));
    pf_batch *b = new pf_batch();
    b->device = device;
    b->ev_ctx.assign(PF_BATCH_MAX, nullptr);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_table, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_snap, hipEventDisableTiming);
    for (int k = 0; k < PF_BATCH_MAX && e == hipSuccess; ++k)
        e = hipEventCreateWithFlags(&b->ev_ctx[k], hipEventDisableTiming);
    void *dt = nullptr, *ht = nullptr;
    if (e == hipSuccess) e = hipMalloc(&dt, kBatchTable);
    if (e == hipSuccess) e = hipHostMalloc(&ht, kBatchTable, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&b->snap, sizeof(pf_state) * PF_BATCH_MAX);
    if (e == hipSuccess) e = hipMalloc((void **)&b->bar, sizeof(pf_gbar));
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->bar_host, sizeof(pf_gbar), hipHostMallocDefault);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->coop, hipDeviceAttributeCooperativeLaunch, device);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->snap_host, sizeof(pf_state) * PF_BATCH_MAX, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventRecord(b->ev_table, b->stream);
    if (dt) {
        b->T = (pf_bslot *)dt;
        b->est = (long *)((char *)dt + sizeof(pf_bslot) * PF_BATCH_MAX);
        b->qst = (int *)((char *)b->est + 2 * sizeof(long) * PF_BATCH_MAX);
    }
    if (ht) {
        b->T_host = (pf_bslot *)ht;
        b->est_host = (long *)((char *)ht + sizeof(pf_bslot) * PF_BATCH_MAX);
        b->qst_host = (int *)((char *)b->est_host + 2 * sizeof(long) * PF_BATCH_MAX);
    }
    if (e != hipSuccess) {
        const int rc = pf_set_hip_error(e, "pf_batch_create", __FILE__, __LINE__);
        pf_batch_destroy(b);
        return rc;
    }
    *out = b;
    return PF_OK;
}

int pf_batch_destroy(pf_batch *b)
{
    if (!b) return PF_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_graph_drop(b);
    if (b->T) (void)hipFree(b->T);
    if (b->T_host) (void)hipHostFree(b->T_host);
    if (b->snap) (void)hipFree(b->snap);
    if (b->snap_host) (void)hipHostFree(b->snap_host);
    if (b->bar) (void)hipFree(b->bar);
    if (b->bar_host) (void)hipHostFree(b->bar_host);
    for (hipEvent_t e : b->ev_ctx)
        if (e) (void)hipEventDestroy(e);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    if (b->ev_table) (void)hipEventDestroy(b->ev_table);
    if (b->ev_snap) (void)hipEventDestroy(b->ev_snap);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PF_OK;
}

int pf_batch_solo_set(pf_batch *b, long min_elems)
{
    if (!b || min_elems < 0) return PF_ERR_ARG;
    if (min_elems != b->solo_elems) {
        b->solo_elems = min_elems;
        b->stale = true;         // the next run rebuilds the table and its graph
    }
    return PF_OK;
}

int pf_batch_run(pf_batch *b, int nctx, pf_ctx *const *ctxs, int n_iters)
{
    if (!b || n_iters < 0) return PF_ERR_ARG;
    int rc = batch_prepare(b, nctx, ctxs);
    if (rc) return rc;
    rc = batch_iterations(b, n_iters);
    if (rc) return rc;
    return batch_release(b, nctx, ctxs);
}

int pf_batch_run_graph(pf_batch *b, int nctx, pf_ctx *const *ctxs, int n_iters)
{
    if (!b || n_iters < 1) return PF_ERR_ARG;
    int rc = batch_prepare(b, nctx, ctxs);
    if (rc) return rc;
    if (!b->gexec || b->graph_iters != n_iters) {
        batch_graph_drop(b);
        PF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = batch_iterations(b, n_iters);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : pf_set_hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
        }
        const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            b->gexec = nullptr;
            return pf_set_hip_error(ei, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        b->graph = g;
        b->graph_iters = n_iters;
    }
    PF_HIP(hipGraphLaunch(b->gexec, b->stream));
    return batch_release(b, nctx, ctxs);
}

int pf_batch_run_persistent(pf_batch *b, int nctx, pf_ctx *const *ctxs, int max_iters)
{
    if (!b || max_iters < 1) return PF_ERR_ARG;
    if (!b->coop) return PF_ERR_ARG;   // the device cannot guarantee a resident grid
    int rc = batch_prepare(b, nctx, ctxs, false);   // every solve in the one kernel
    if (rc) return rc;
    switch (b->ntot) {
    case 60: rc = batch_persistent_nt<60>(b, max_iters); break;
    case 24: rc = batch_persistent_nt<24>(b, max_iters); break;
    default: rc = batch_persistent_nt<8>(b, max_iters); break;
    }
    if (rc) return rc;
    b->persistent_used = true;
    return batch_release(b, nctx, ctxs);
}

int pf_batch_query(pf_batch *b, pf_pcg_result *res, int *ready)
{
    if (!b || !res || !ready) return PF_ERR_ARG;
    if (b->ns < 1) return PF_ERR_STATE;
    *ready = 0;
    if (!b->snap_pending) {
        hipLaunchKernelGGL(kb_snapshot, dim3(b->ns), dim3(64), 0, b->stream, b->T, b->ns, b->snap);
        PF_HIP(hipGetLastError());
        PF_HIP(hipMemcpyAsync(b->snap_host, b->snap, sizeof(pf_state) * b->ns, hipMemcpyDeviceToHost, b->stream));
        if (b->persistent_used)
            PF_HIP(hipMemcpyAsync(b->bar_host, b->bar, sizeof(pf_gbar), hipMemcpyDeviceToHost, b->stream));
        PF_HIP(hipEventRecord(b->ev_snap, b->stream));
        b->snap_pending = true;
        b->snap_n = b->ns;
    }
    const hipError_t q = hipEventQuery(b->ev_snap);
    if (q == hipErrorNotReady) return PF_OK;
    if (q != hipSuccess) return pf_set_hip_error(q, "hipEventQuery", __FILE__, __LINE__);
    b->snap_pending = false;
    if (b->persistent_used && b->bar_host->error) {
        // a persistent run's grid barrier ran out of time: its solves stopped
        // part-way through an iteration
        return pf_set_hip_error(hipErrorLaunchTimeOut, "persistent batch grid barrier", __FILE__, __LINE__);
    }
    for (int k = 0; k < b->snap_n; ++k) fill_result(b->snap_host[k], &res[k]);
    *ready = 1;
    return PF_OK;
}

}  // extern "C"
