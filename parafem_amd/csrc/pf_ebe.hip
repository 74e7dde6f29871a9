// Element-by-element PCG on the device: gather -> element products ->
// scatter, the FP64 reductions, the vector recurrences and the stopping rule.
//
// Reference flow: src/programs/5th_ed/p121/p121.f90:86-104 and
// src/programs/5th_ed/p123/p123.f90:120-151; GATHER/SCATTER/SCATTER_NOADD
// src/modules/mpi/gather_scatter.f90:547-993; DOT_PRODUCT_P/SUM_P
// src/modules/mpi/maths.f90:168-216,271-315; checon_par maths.f90:999-1069.
//
// Per PCG iteration four kernels run on the context's stream:
//   k_ebe      gathers p through the steering (p has a leading 0 so a
//              restrained freedom reads 0), forms u_e = K_e p_e with one
//              64-lane wavefront per hex20 element (lane = row, K read one
//              column at a time, p_e broadcast from lane j), writes utemp;
//   k_scatter  sums each equation's element slots in ascending element order
//              (the order SCATTER adds them at one rank), applies the p123
//              fixed-freedom rows, accumulates p.u and, in its last block,
//              alpha = up / (p.u);
//   k_update   xnew = x + p*alpha, r = r - u*alpha, d = dinv*r, partial
//              (r,d), (r,r), max|xnew|, max|xnew-x|; last block: beta,
//              stopping rule, iteration count, freeze flag;
//   k_pupdate  p = d + p*beta.
// Every kernel returns at once when the freeze flag is set, so iterations
// enqueued after convergence are no-ops and the count equals the host loop's.
//
// All FP64 vector arithmetic is written with explicit roundings (__dadd_rn,
// __dmul_rn) so the compiler cannot fuse it; the element product is a chain
// of fused multiply-adds in ascending column order.  Reductions use one fixed
// order (pf_reduce_grid in the public header), so results are reproducible
// bit for bit from run to run.
#include <cmath>
#include <vector>

#include "pf_internal.h"

namespace {

constexpr int MAXG = PF_REDUCE_MAX_GRID;

__device__ __forceinline__ double bcast64(double v, int src)
{
    // v_readlane on both halves: src is a compile-time constant in the
    // unrolled loops below, so this is two scalar reads and no LDS traffic.
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffu), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
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
// stores the block results in partials[q*MAXG + blockIdx.x]; returns true in
// every thread of the block that finished last.
template <int NQ>
__device__ bool block_to_partials(double (&v)[NQ], const int (&qid)[NQ], const bool (&ismax)[NQ],
                                  double *partials, unsigned int *ticket)
{
    __shared__ double ws[NQ][PF_BLOCK / 64];
    __shared__ int last;
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
            partials[qid[k] * MAXG + blockIdx.x] = s;
        }
        __threadfence();
        const unsigned int t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    return last != 0;
}

// Run by the last block only: folds the G block results of each quantity in
// the fixed order and returns them in thread 0 (out[] valid there).
template <int NQ>
__device__ void fold_partials(double (&out)[NQ], const int (&qid)[NQ], const bool (&ismax)[NQ],
                              const double *partials, unsigned int *ticket)
{
    __shared__ double ws[NQ][PF_BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __threadfence();
    const int G = gridDim.x;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < G; b += PF_BLOCK) {
            const double pv = partials[qid[k] * MAXG + b];
            s = ismax[k] ? fmax(s, pv) : __dadd_rn(s, pv);
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
        *ticket = 0u;
    }
}

// Width of the lane group serving one element, and elements per wavefront.
template <int NT> struct Lanes {
    static constexpr int W = NT > 32 ? 64 : (NT > 16 ? 32 : (NT > 8 ? 16 : 8));
    static constexpr int EPW = 64 / W;
    static constexpr int EPB = EPW * (PF_BLOCK / 64);   // elements per block
};

// Element products in FP64.  GATHER: p is the neq+1 vector with p[0] = 0 and
// the element vector is p[g_g]; otherwise src is pmul (ntot,nels).
template <int NT, bool GATHER>
__global__ __launch_bounds__(PF_BLOCK) void k_ebe_f64(const double *__restrict__ K, const int *__restrict__ gg,
                                                     const double *__restrict__ src, double *__restrict__ utemp,
                                                     int nels, const pf_state *__restrict__ st)
{
    if (st && st->frozen) return;
    using L = Lanes<NT>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sub = lane / L::W, li = lane % L::W;
    const long e = ((long)blockIdx.x * (PF_BLOCK / 64) + wave) * L::EPW + sub;
    const long ec = e < nels ? e : (long)nels - 1;
    const int i = li < NT ? li : NT - 1;
    const double *__restrict__ Ke = K + ec * (long)(NT * NT) + i;
    // Issue every column load of K_e before the dependent steering -> p
    // gather so the element's whole matrix is in flight at once.
    double kc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) kc[j] = __builtin_nontemporal_load(&Ke[j * NT]);
    __builtin_amdgcn_sched_barrier(0);
    const double pe = GATHER ? src[gg[ec * NT + i]] : src[ec * NT + i];
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const double pj = (L::W == 64) ? bcast64(pe, j) : __shfl(pe, sub * L::W + j, 64);
        acc = __fma_rn(kc[j], pj, acc);
    }
    if (e < nels && li < NT) utemp[e * NT + li] = acc;
}

// SCATTER fused with p.u.  u and p have the leading 0 slot.
__global__ __launch_bounds__(PF_BLOCK) void k_scatter_pu(const int *__restrict__ ptr, const int *__restrict__ slot,
                                                        const double *__restrict__ utemp,
                                                        const double *__restrict__ fixstore,
                                                        const double *__restrict__ p, double *__restrict__ u,
                                                        int neq, double *partials, pf_state *st)
{
    if (st->frozen) return;
    double acc[1] = {0.0};
    const int stride = gridDim.x * PF_BLOCK;
    for (int q = blockIdx.x * PF_BLOCK + threadIdx.x; q < neq; q += stride) {
        const int a = ptr[q], b = ptr[q + 1];
        double s = 0.0;
        for (int k = a; k < b; ++k) s = __dadd_rn(s, utemp[slot[k]]);
        const double pq = p[q + 1];
        if (fixstore) {
            const double fs = fixstore[q + 1];
            if (fs != 0.0) s = __dmul_rn(pq, fs);
        }
        u[q + 1] = s;
        acc[0] = __dadd_rn(acc[0], __dmul_rn(pq, s));
    }
    constexpr int qid[1] = {PF_Q_PU};
    constexpr bool mx[1] = {false};
    if (block_to_partials<1>(acc, qid, mx, partials, &st->ticket[0])) {
        double out[1];
        fold_partials<1>(out, qid, mx, partials, &st->ticket[0]);
        if (threadIdx.x == 0) {
            const double alpha = st->up / out[0];
            st->alpha = alpha;
            if (!isfinite(alpha)) {
                st->breakdown = 1;
                st->frozen = 1;
            }
        }
    }
}

// x/r/d updates, (r,d), (r,r), the checon maxima, then beta and the rule.
__global__ __launch_bounds__(PF_BLOCK) void k_update(double *__restrict__ x, double *__restrict__ r,
                                                    double *__restrict__ d, const double *__restrict__ p,
                                                    const double *__restrict__ u, const double *__restrict__ dinv,
                                                    int neq, double *partials, pf_state *st)
{
    if (st->frozen) return;
    const double alpha = st->alpha;
    double v[4] = {0.0, 0.0, 0.0, 0.0};   // rd, rr, max|xnew|, max|xnew-x|
    const int stride = gridDim.x * PF_BLOCK;
    for (int q = blockIdx.x * PF_BLOCK + threadIdx.x + 1; q <= neq; q += stride) {
        const double xo = x[q];
        const double xn = __dadd_rn(xo, __dmul_rn(p[q], alpha));
        const double rq = __dsub_rn(r[q], __dmul_rn(u[q], alpha));
        const double dq = __dmul_rn(dinv[q], rq);
        x[q] = xn;
        r[q] = rq;
        d[q] = dq;
        v[0] = __dadd_rn(v[0], __dmul_rn(rq, dq));
        v[1] = __dadd_rn(v[1], __dmul_rn(rq, rq));
        v[2] = fmax(v[2], fabs(xn));
        v[3] = fmax(v[3], fabs(__dsub_rn(xn, xo)));
    }
    constexpr int qid[4] = {PF_Q_RD, PF_Q_RR, PF_Q_MAXX, PF_Q_MAXDX};
    constexpr bool mx[4] = {false, false, true, true};
    if (block_to_partials<4>(v, qid, mx, partials, &st->ticket[1])) {
        double out[4];
        fold_partials<4>(out, qid, mx, partials, &st->ticket[1]);
        if (threadIdx.x == 0) {
            const double rd = out[0], rr = out[1];
            const double beta = rd / st->up;
            const double crit = st->stop_mode == PF_STOP_CHECON ? out[3] / out[2] : sqrt(rr / st->rr0);
            st->beta = beta;
            st->up = rd;
            st->rr = rr;
            st->crit = crit;
            st->iters += 1;
            st->converged = crit <= st->tol;
            if (!isfinite(beta) || !isfinite(rd)) st->breakdown = 1;
            if (st->converged || st->iters >= st->limit || st->breakdown) st->frozen = 1;
        }
    }
}

__global__ __launch_bounds__(PF_BLOCK) void k_pupdate(double *__restrict__ p, const double *__restrict__ d, int neq,
                                                     const pf_state *__restrict__ st)
{
    if (st->frozen) return;
    const double beta = st->beta;
    const int stride = gridDim.x * PF_BLOCK;
    for (int q = blockIdx.x * PF_BLOCK + threadIdx.x + 1; q <= neq; q += stride)
        p[q] = __dadd_rn(d[q], __dmul_rn(p[q], beta));
}

// PCG start: dinv = 1/diag, d = dinv*r, p = d, x = 0, up = (r,d), rr0 = (r,r).
__global__ __launch_bounds__(PF_BLOCK) void k_pcg_init(double *__restrict__ x, const double *__restrict__ r,
                                                      double *__restrict__ d, double *__restrict__ p,
                                                      double *__restrict__ dinv, const double *__restrict__ diag,
                                                      int neq, double *partials, pf_state *st)
{
    double v[2] = {0.0, 0.0};
    const int stride = gridDim.x * PF_BLOCK;
    for (int q = blockIdx.x * PF_BLOCK + threadIdx.x + 1; q <= neq; q += stride) {
        const double di = 1.0 / diag[q];
        const double rq = r[q];
        const double dq = __dmul_rn(di, rq);
        dinv[q] = di;
        d[q] = dq;
        p[q] = dq;
        x[q] = 0.0;
        v[0] = __dadd_rn(v[0], __dmul_rn(rq, dq));
        v[1] = __dadd_rn(v[1], __dmul_rn(rq, rq));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        x[0] = 0.0; d[0] = 0.0; p[0] = 0.0; dinv[0] = 0.0;
    }
    constexpr int qid[2] = {PF_Q_RD, PF_Q_RR};
    constexpr bool mx[2] = {false, false};
    if (block_to_partials<2>(v, qid, mx, partials, &st->ticket[2])) {
        double out[2];
        fold_partials<2>(out, qid, mx, partials, &st->ticket[2]);
        if (threadIdx.x == 0) {
            st->up = out[0];
            st->rr0 = out[1];
            st->rr = out[1];
            st->iters = 0;
            st->converged = 0;
            st->breakdown = !isfinite(out[0]) || !isfinite(out[1]);
            // A zero right-hand side is already solved by x = 0.
            st->frozen = (out[1] == 0.0) || st->breakdown;
            st->converged = out[1] == 0.0;
            st->crit = 0.0;
        }
    }
}

// ---------------------------------------------------------- primitives
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

// Generic reduction for the primitives: mode 0 = sum a*b, 1 = sum a,
// 2 = max|a| and max|a-b| then b = a (checon_par).  Result in out[0..1].
__global__ __launch_bounds__(PF_BLOCK) void k_reduce_prim(const double *__restrict__ a, double *__restrict__ b,
                                                         long n, int mode, double *partials, unsigned int *ticket,
                                                         double *out)
{
    double v[2] = {0.0, 0.0};
    const long stride = (long)gridDim.x * PF_BLOCK;
    for (long i = (long)blockIdx.x * PF_BLOCK + threadIdx.x; i < n; i += stride) {
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
    if (mode == 2) {
        constexpr int qid[2] = {0, 1};
        constexpr bool mx[2] = {true, true};
        if (block_to_partials<2>(v, qid, mx, partials, ticket)) {
            double o[2];
            fold_partials<2>(o, qid, mx, partials, ticket);
            if (threadIdx.x == 0) { out[0] = o[0]; out[1] = o[1]; }
        }
    } else {
        double w[1] = {v[0]};
        constexpr int qid[1] = {0};
        constexpr bool mx[1] = {false};
        if (block_to_partials<1>(w, qid, mx, partials, ticket)) {
            double o[1];
            fold_partials<1>(o, qid, mx, partials, ticket);
            if (threadIdx.x == 0) out[0] = o[0];
        }
    }
}

// y = K x in FP64 with the fixed rows (pf_apply): scatter into y (no ext).
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

template <bool GATHER>
int launch_ebe_f64(pf_ctx *c, const double *src, double *utemp, const pf_state *st)
{
    int grid;
    switch (c->ntot) {
    case 60:
        grid = (c->nels + Lanes<60>::EPB - 1) / Lanes<60>::EPB;
        hipLaunchKernelGGL((k_ebe_f64<60, GATHER>), dim3(grid), dim3(PF_BLOCK), 0, c->stream, c->K64, c->g_g, src,
                           utemp, c->nels, st);
        break;
    case 24:
        grid = (c->nels + Lanes<24>::EPB - 1) / Lanes<24>::EPB;
        hipLaunchKernelGGL((k_ebe_f64<24, GATHER>), dim3(grid), dim3(PF_BLOCK), 0, c->stream, c->K64, c->g_g, src,
                           utemp, c->nels, st);
        break;
    case 8:
        grid = (c->nels + Lanes<8>::EPB - 1) / Lanes<8>::EPB;
        hipLaunchKernelGGL((k_ebe_f64<8, GATHER>), dim3(grid), dim3(PF_BLOCK), 0, c->stream, c->K64, c->g_g, src,
                           utemp, c->nels, st);
        break;
    default:
        return PF_ERR_ARG;
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int reduce_prim(pf_ctx *c, const double *a, double *b, int64_t n, int mode, double *out2)
{
    if (n <= 0) {
        out2[0] = out2[1] = 0.0;
        return PF_OK;
    }
    const int G = pf_grid_for(n);
    double *dout = c->partials + (size_t)PF_NQ * MAXG;   // 2 spare slots past the PCG partials
    hipLaunchKernelGGL(k_reduce_prim, dim3(G), dim3(PF_BLOCK), 0, c->stream, a, b, (long)n, mode, c->partials,
                       &c->state->ticket[3], dout);
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(c->scalar_host, dout, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    out2[0] = c->scalar_host[0];
    out2[1] = c->scalar_host[1];
    return PF_OK;
}

}  // namespace

extern "C" {

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
    return launch_ebe_f64<false>(c, pmul, utemp, nullptr);
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

int pf_checon_par(pf_ctx *c, const double *xnew, double tol, int *converged, double *x, int64_t n)
{
    if (!c || !converged || n < 0 || (n && (!xnew || !x))) return PF_ERR_ARG;
    double o[2];
    int rc = reduce_prim(c, xnew, x, n, 2, o);
    if (!rc) *converged = (o[1] / o[0]) <= tol;
    return rc;
}

int pf_pcg_begin(pf_ctx *c, const double *r0, int r0_on_device, const pf_pcg_opts *opts)
{
    if (!c || !r0 || !opts) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    if (opts->limit < 1 || !(opts->tol >= 0.0)) return PF_ERR_ARG;
    if (opts->stop_mode != PF_STOP_CHECON && opts->stop_mode != PF_STOP_RELRES) return PF_ERR_ARG;
    if (opts->precision != PF_PREC_FP64) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    c->opts = *opts;
    PF_HIP(hipMemcpyAsync(c->r + 1, r0, sizeof(double) * c->neq,
                          r0_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    pf_state s{};
    s.tol = opts->tol;
    s.limit = opts->limit;
    s.stop_mode = opts->stop_mode;
    *c->state_host = s;
    PF_HIP(hipMemcpyAsync(c->state, c->state_host, sizeof(pf_state), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_pcg_init, dim3(pf_grid_for(c->neq)), dim3(PF_BLOCK), 0, c->stream, c->x, c->r, c->d, c->p,
                       c->dinv, c->diag, c->neq, c->partials, c->state);
    PF_HIP(hipGetLastError());
    // The host mirror is read back by pf_pcg_poll; wait so it is not reused
    // while the copy above is still in flight.
    PF_HIP(hipStreamSynchronize(c->stream));
    c->pcg_ready = true;
    return PF_OK;
}

int pf_pcg_run(pf_ctx *c, int n)
{
    if (!c || n < 0) return PF_ERR_ARG;
    if (!c->pcg_ready) return PF_ERR_STATE;
    const int G = pf_grid_for(c->neq);
    hipEvent_t t0, t1;
    for (int it = 0; it < n; ++it) {
        bool tm = pf_timing_begin(c, PF_K_EBE, &t0, &t1);
        if (tm) PF_HIP(hipEventRecord(t0, c->stream));
        int rc = launch_ebe_f64<true>(c, c->p, c->utemp, c->state);
        if (rc) return rc;
        if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        tm = pf_timing_begin(c, PF_K_SCATTER, &t0, &t1);
        if (tm) PF_HIP(hipEventRecord(t0, c->stream));
        hipLaunchKernelGGL(k_scatter_pu, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->csr_ptr, c->csr_slot, c->utemp,
                           c->fixstore, c->p, c->u, c->neq, c->partials, c->state);
        if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        tm = pf_timing_begin(c, PF_K_UPDATE, &t0, &t1);
        if (tm) PF_HIP(hipEventRecord(t0, c->stream));
        hipLaunchKernelGGL(k_update, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->x, c->r, c->d, c->p, c->u, c->dinv,
                           c->neq, c->partials, c->state);
        if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        tm = pf_timing_begin(c, PF_K_PUPDATE, &t0, &t1);
        if (tm) PF_HIP(hipEventRecord(t0, c->stream));
        hipLaunchKernelGGL(k_pupdate, dim3(G), dim3(PF_BLOCK), 0, c->stream, c->p, c->d, c->neq, c->state);
        if (tm) PF_HIP(hipEventRecord(t1, c->stream));
        PF_HIP(hipGetLastError());
    }
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

int pf_pcg_timing_get(pf_ctx *c, double ms[4], int counts[4])
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
    const pf_state &s = *c->state_host;
    res->iters = s.iters;
    res->converged = s.converged;
    res->crit = s.crit;
    res->rr0 = s.rr0;
    res->rr = s.rr;
    res->refinements = 0;
    res->status = s.breakdown ? PF_ERR_BREAKDOWN : (s.converged ? PF_OK : PF_NOT_CONVERGED);
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

int pf_pcg_solve(pf_ctx *c, const double *r0, double *x, const pf_pcg_opts *opts, pf_pcg_result *res)
{
    if (!c || !r0 || !x || !opts) return PF_ERR_ARG;
    int rc = pf_pcg_begin(c, r0, 0, opts);
    if (rc) return rc;
    pf_pcg_result tmp{};
    pf_pcg_result *R = res ? res : &tmp;
    const int chunk = opts->check_every > 0 ? opts->check_every : 16;
    rc = pf_pcg_poll(c, R);
    while (!rc && !c->state_host->frozen) {
        int n = chunk;
        const int left = opts->limit - R->iters;
        if (n > left) n = left;
        if (n < 1) n = 1;
        rc = pf_pcg_run(c, n);
        if (!rc) rc = pf_pcg_poll(c, R);
    }
    if (rc) return rc;
    rc = pf_pcg_x(c, x, 0);
    if (rc) return rc;
    return R->status;
}

int pf_apply(pf_ctx *c, const double *x_host, double *y_host)
{
    if (!c || !x_host || !y_host) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
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
    if (!rc) rc = launch_ebe_f64<true>(c, xe, ut, nullptr);
    if (!rc) {
        hipLaunchKernelGGL(k_scatter_fix, dim3(pf_grid_for(c->neq)), dim3(256), 0, c->stream, c->csr_ptr,
                           c->csr_slot, ut, c->fixstore, xe, y, c->neq);
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

}  // extern "C"
