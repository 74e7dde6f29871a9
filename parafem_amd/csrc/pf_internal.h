// Internal declarations shared by the HIP translation units of libparafem_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/parafem_amd.h"

#define PF_BLOCK 256                 // threads per block of the vector kernels
#define PF_REDUCE_MAX_GRID 16384     // upper bound on reduction grid size G

// Kernels of one PCG iteration, in launch order (pf_pcg_timing_get indices):
// element products, scatter + p.u (and its fold), x/r/d update (and its
// fold), p update, and the mixed solver's true-residual replacement (all of
// its launches, which are no-ops on iterations that do not replace r).
enum { PF_K_EBE = 0, PF_K_SCATTER = 1, PF_K_UPDATE = 2, PF_K_PUPDATE = 3, PF_K_REFINE = 4, PF_NKERNELS = 5 };

// Per-block partial quantities kept in the partials buffer.
enum { PF_Q_PU = 0, PF_Q_RD = 1, PF_Q_RR = 2, PF_Q_MAXX = 3, PF_Q_MAXDX = 4, PF_NQ = 5 };

// Internal element-matrix layouts (the C ABI speaks storkm_pp's column-major
// order; pf_storkm_set/get convert).  For element e, row i, column j of an
// ntot x ntot matrix:
//   FP64 "column pairs": e*ntot^2 + (j/2)*2*ntot + 2*i + (j%2)
//     -> lane i reads columns (j, j+1) as one 16-byte load
//   FP32 hi/lo: packed symmetric (below).
__host__ __device__ inline long pf_idx_pair(long e, int nt, int i, int j)
{
    return e * (long)nt * nt + (long)(j >> 1) * 2 * nt + 2 * i + (j & 1);
}

// Packed symmetric storage of the FP32 hi and lo parts.  Row i uses the
// nt/2 + 1 entries K(i, (i+d) mod nt), d = 0..nt/2, taken from the upper
// triangle K(min, max); the entries of row i left of the diagonal band are
// the band entries of other rows (K(i, i-d) = K(i-d, (i-d)+d)).  Per element:
// the nt/2 full diagonals d < nt/2 in quads, lane-major (16-byte loads),
// then their remainder (nt/2 mod 4, lane-major), then the last diagonal
// d = nt/2 once: its nt/2 distinct entries, row i reading entry i mod nt/2
// (K(i, i+nt/2) = K(i+nt/2, i)).  nt(nt+1)/2 values, padded to a multiple of
// 4 so that every element starts 16-byte aligned.
__host__ __device__ constexpr int pf_sym_diags(int nt) { return nt / 2 + 1; }
__host__ __device__ constexpr int pf_sym_quads(int nt) { return (nt / 2) / 4; }
__host__ __device__ constexpr int pf_sym_rem(int nt) { return nt / 2 - 4 * pf_sym_quads(nt); }
__host__ __device__ constexpr long pf_sym_per(int nt) { return ((long)nt * (nt + 1) / 2 + 3) & ~3L; }
__host__ __device__ inline long pf_idx_sym(long e, int nt, int i, int d)
{
    const int Q4 = 4 * pf_sym_quads(nt), RF = pf_sym_rem(nt), H = nt / 2;
    const long base = e * pf_sym_per(nt);
    if (d < Q4) return base + (long)(d >> 2) * 4 * nt + 4 * i + (d & 3);
    if (d < H) return base + (long)Q4 * nt + (long)RF * i + (d - Q4);
    return base + (long)H * nt + i % H;
}
// Packed slot of K(i, j).
__host__ __device__ inline long pf_idx_sym_ij(long e, int nt, int i, int j)
{
    int d = j - i;
    if (d < 0) d += nt;
    return d <= nt / 2 ? pf_idx_sym(e, nt, i, d) : pf_idx_sym(e, nt, j, nt - d);
}

// Device-resident PCG scalars.  Written only by the one-block fold kernels;
// read at the start of the following kernels.
struct pf_state {
    double up;        // (r,d) of the current iterate (the reference's "up")
    double alpha;
    double beta;
    double rr0;       // ||r0||^2
    double rr;        // ||r||^2 (recurrence residual, or true residual after a replacement)
    double rr_ref;    // ||r||^2 at the last true-residual replacement (mixed)
    double crit;      // last stopping criterion value
    double tol;
    double delta2;    // mixed: replacement trigger (delta^2)
    int    iters;
    int    limit;
    int    stop_mode;
    int    precision;
    int    converged;
    int    frozen;    // 1 once converged, at the limit, or broken down
    int    breakdown;
    int    refine;    // mixed: a true-residual replacement is due this iteration
    int    refinements;
    int    refine_every;
    int    since_refine;
    int    pfresh;    // p holds this iteration's direction (the start); otherwise the
                      // range-fused products of hex8 meshes form p = d + beta p
};

// Multi-RHS block solve (pf_block_solve): up to PF_BLK right-hand sides as
// the columns of row-major (neq+1) x PF_BLK block vectors.  Column c is kept
// at position pf_blk_pos(c) so that lane group (c % 16) of an MFMA tile reads
// its four columns c % 16 + 16 nt, nt = 0..3, as one vector.
#define PF_BLK 64
#define PF_BLK_MAXG 2048
__host__ __device__ inline int pf_blk_pos(int c) { return (c & 15) * 4 + (c >> 4); }

struct pf_blk_state {
    double up[PF_BLK], alpha[PF_BLK], beta[PF_BLK], rr[PF_BLK];
    double rr_ref[PF_BLK];   // ||r||^2 at the column's last residual replacement
    double ff[PF_BLK];       // ||f||^2
    int frozen[PF_BLK];      // converged (true residual) or unused column
    int iters, limit, done, trig;   // trig: some column asks for a replacement
    int xlast;               // done was set this iteration: its p update still adds alpha p into x
    double tol2, delta2;
};

struct pf_block {
    float *Kt32 = nullptr;     // element matrices as FP32 MFMA A-operand tiles, upper tile triangle (nels * 2560 for hex20)
    double *Kt64 = nullptr;    // the same in FP64
    float *P = nullptr, *U = nullptr, *R = nullptr, *X = nullptr;   // FP32 inner vectors
    float *dinv = nullptr;     // (neq+1) Jacobi, FP32
    float *Ut32 = nullptr;     // (nels*ntot) x PF_BLK element products, FP32
    double *Ut64 = nullptr;    // the same in FP64 (true residual)
    double *X64 = nullptr, *R64 = nullptr, *F64 = nullptr, *U64 = nullptr;
    double *stage = nullptr;   // neq x PF_BLK host-order staging (upload/download)
    double *partials = nullptr;   // 3 x PF_BLK_MAXG x PF_BLK
    pf_blk_state *st = nullptr, *st_host = nullptr;
};

// Range-fused products and sums of the mixed solver (pf_fused.hip): ranges
// of consecutive elements, each range's distinct equations (with first-writer
// and fixed-row flags), the local index of every element slot, the rounds
// inside a range and the ranges grouped by colour.
struct pf_ranges {
    unsigned short *lmap = nullptr;    // (ntot,nels) local index in the range's equation list, 0xFFFF restrained
    unsigned char *lround = nullptr;   // (nels) round of the element inside its range
    int *rptr = nullptr;               // (nrange+1) offsets into ue
    unsigned *ue = nullptr;            // distinct equations per range, ascending, flag bits on top
    int *cperm = nullptr;              // ranges by colour
    double *rpart = nullptr;           // (nrange) p.u per range
    long nslot = 0, nels = 0, nrange = 0, nue = 0;
    int nround = 1, ncol = 1;
    std::vector<int> cstart;           // (ncol+1) colour offsets into cperm
};

struct pf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;      // stream in use
    hipStream_t own_stream = nullptr;  // created by the context
    size_t bytes = 0;                  // device bytes in live allocations
    // Freed device blocks kept for reuse by later allocations of this
    // context (hipFree waits for the device; a context that solves one
    // problem after another reuses its buffers instead).
    std::vector<std::pair<void *, size_t>> cache;
    std::vector<std::pair<void *, size_t>> live;   // allocation -> block size
    size_t cached = 0;

    // mesh
    int nod = 0, nels = 0;
    double *g_coord = nullptr;         // (nod,3,nels)
    // steering
    int ntot = 0, neq = 0;
    int *g_g = nullptr;                // (ntot,nels), 1-based, 0 = restrained
    int *csr_ptr = nullptr;            // (neq+1) offsets into csr_slot
    int *csr_slot = nullptr;           // element slots e*ntot+i, ascending
    int *csr_slot_t = nullptr;         // hex8: the same list as i*nels+e (row-major iteration products)
    int64_t nnz = 0;
    // element matrices (internal layouts above)
    int store = PF_STORE_FP64;
    double *K64 = nullptr;             // FP64 column pairs (PF_STORE_FP64)
    double *K64b = nullptr;            // p124's second matrix set (storkb_pp), FP64 column pairs
    bool k_shared = false;             // K64 holds one matrix used by every element
    float *K32hi = nullptr;            // packed symmetric FP32, hi = fl32(K)
    float *K32lo = nullptr;            // packed symmetric FP32, lo = fl32(K - hi)
    bool have_K = false;
    // p123 fixed freedoms: fixstore[eq] = store value, 0 for free equations
    double *fixstore = nullptr;
    int nfix = 0;
    bool fix_zero = false;             // fixed rows give u = 0 instead of p*store (p122, later plastic iterations)

    // PCG vectors: length neq+1, entry 0 is a permanent zero so that a
    // steering value of 0 gathers 0 without a branch.
    double *x = nullptr, *r = nullptr, *p = nullptr, *d = nullptr, *u = nullptr;
    double *dinv = nullptr, *diag = nullptr, *f = nullptr;
    double *diag0 = nullptr;           // element diagonal before penalty rows
    float *p32 = nullptr;              // fl32(p): the FP32 products' gather source (mixed path)
    double *utemp = nullptr;           // (ntot,nels) element products, FP64
    float *utemp32 = nullptr;          // (ntot,nels) element products, FP32
    double *partials = nullptr;        // PF_NQ * PF_REDUCE_MAX_GRID (+ spare)
    pf_state *state = nullptr;         // device
    pf_state *state_host = nullptr;    // pinned mirror for polling
    double *scalar_host = nullptr;     // pinned scratch for primitive results
    pf_pcg_opts opts{};
    bool pcg_ready = false;
    bool precon_given = false;         // dinv set by pf_precon_set, not 1/diag
    pf_block *blk = nullptr;           // multi-RHS buffers (allocated on first use)
    // Non-blocking state read-back (pf_pcg_query).
    pf_state *state_query = nullptr;   // pinned
    hipEvent_t ev_query = nullptr;
    bool query_pending = false;
    // Captured iteration chunk (pf_pcg_run_graph) and what it was captured for.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_iters = 0, graph_precision = -1, graph_stop = -1, graph_dot = -1;
    // Summation order of dots and sums (PF_DOT_TREE or PF_DOT_SEQ).
    int dot_order = PF_DOT_TREE;
    // How the mixed solver sums element products into u (PF_SCATTER_*), and
    // the ranges of the range-fused mode (built at pf_pcg_begin).
    int scatter_mode = PF_SCATTER_SLOTS;
    int graph_scatter = -1;
    pf_ranges *rng = nullptr;
    // Launch shapes (pf_tune_set); 0 = the built-in choice.
    int tune_epl = 0;                  // single-RHS products: elements per lane group
    int tune_blk_waves = 0;            // block products: elements per workgroup
    int tune_blk_rows = 0;             // block FP64 products: 16-row tiles per pass

    // Optional per-kernel timing of pf_pcg_run: one event pair per launch.
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_used[PF_NKERNELS];   // (start, stop) indices
    size_t ev_next = 0;
};

// Range-fused mixed iteration (pf_fused.hip): build / free the ranges of the
// current steering and fixed rows; enqueue the products with their sums into
// u and the p.u partials; fold alpha from them.
int pf_ranges_build(pf_ctx *c);
void pf_ranges_drop(pf_ctx *c);
int pf_ranges_iteration(pf_ctx *c);
int pf_ranges_fold(pf_ctx *c);
bool pf_ranges_form_p(const pf_ctx *c);   // the range kernels form p (no p pass)

// Exclusive scan of n integer counts on the context's stream (pf_setup.hip):
// ptr[q + 1] = cnt[0] + ... + cnt[q] (ptr[0] is the caller's), the sum into
// *total, bit 2 into *zero_flag where a count is zero; bsum holds
// pf_scan_blocks(n) ints.
int pf_scan_blocks(long n);
void pf_scan_launch(pf_ctx *c, const int *cnt, int n, int *bsum, int *ptr, int *total, int *zero_flag);

// Looks up, once per device (the device current), the address of the
// allocation-free entry points' scratch array (pf_ebe.hip).
int pf_scratch_init(int device);

// p124's transient element matrices into K64 (storka) and K64b (storkb).
int pf_launch_assemble_transient(pf_ctx *c, double theta, double dtim, const double *prop, const int *etype);

// Steering set-up (pf_context.hip, pf_setup.hip).
int pf_steer_install(pf_ctx *c, int ntot, int nels, int neq);
void pf_steer_drop(pf_ctx *c);
int pf_csr_build(pf_ctx *c);

// Largest elements-per-lane-group of the single-RHS element product.
inline int pf_epl_max(int ntot) { return ntot > 32 ? 2 : ntot > 8 ? 4 : 8; }

// Drop a captured iteration chunk (problem, options or buffers changed).
void pf_graph_drop(pf_ctx *c);
// Free the multi-RHS buffers (problem or matrices changed).
void pf_block_free(pf_ctx *c);

// Event pair handles for timing kernel `k` (returns false when timing is off).
bool pf_timing_begin(pf_ctx *c, int k, hipEvent_t *start, hipEvent_t *stop);

// error helpers -------------------------------------------------------------
#define PF_HIP(call)                                        \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return pf_set_hip_error(_e, #call, __FILE__, __LINE__); \
    } while (0)

// FP64 element-matrix entries held (one matrix when shared) and the element
// stride of the kernels' K64 indexing (0 when shared).
inline size_t pf_k64_count(const pf_ctx *c)
{
    return (size_t)c->ntot * c->ntot * (c->k_shared ? 1 : (size_t)c->nels);
}
inline long pf_k64_estride(const pf_ctx *c) { return c->k_shared ? 0 : 1; }

int pf_set_hip_error(hipError_t e, const char *what, const char *file, int line);
int pf_alloc(pf_ctx *c, void **p, size_t bytes);
void pf_free(pf_ctx *c, void *p, size_t bytes);
void pf_cache_release(pf_ctx *c);
int pf_grid_for(int64_t n);

// kernels launched across translation units ---------------------------------
// Element assembly into K64 (FP64 storage) or, for FP32 storage, straight
// into K32hi/K32lo with the FP64 element diagonals in dg (ntot*nels).
int pf_launch_assemble_elastic(pf_ctx *c, double v2, double vv, double den, double e0, const double *e_elem_dev,
                               double *dg);
int pf_launch_assemble_laplace(pf_ctx *c, double kx, double ky, double kz, const double *k_elem_dev, double *dg);
// Assembled diagonal (leading 0 slot) from element diagonals dg(ntot,nels).
int pf_launch_diag_elem(pf_ctx *c, const double *dg, double *diag);
int pf_launch_split_fp32(pf_ctx *c);
int pf_launch_diag(pf_ctx *c, double *diag_dev);
// storkm_pp (column-major, nels_chunk elements starting at e0) <-> internal
int pf_launch_from_colmajor(pf_ctx *c, const double *src, long e0, long nel);
int pf_launch_to_colmajor(pf_ctx *c, double *dst, long e0, long nel);
