// Context, device memory and set-up entry points of libparafem_amd.
//
// One pf_ctx drives one GPU from one host thread.  It owns every device
// buffer of a problem: element coordinates, steering, the transposed
// steering used by the deterministic scatter, element matrices and the PCG
// vectors.  Host arrays passed in are only read during the call.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pf_internal.h"

static thread_local char g_last_error[512];

int pf_set_hip_error(hipError_t e, const char *what, const char *file, int line)
{
    snprintf(g_last_error, sizeof g_last_error, "%s failed: %s (%s:%d)", what,
             hipGetErrorString(e), file, line);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return PF_ERR_OOM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PF_ERR_NODEVICE;
    return PF_ERR_HIP;
}

int pf_alloc(pf_ctx *c, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    // best fit among the cached blocks no larger than 2x the request + 1 MiB
    size_t best = (size_t)-1, pick = 0;
    for (size_t k = 0; k < c->cache.size(); ++k) {
        const size_t sz = c->cache[k].second;
        if (sz >= bytes && sz <= 2 * bytes + (1u << 20) && sz < best) {
            best = sz;
            pick = k;
        }
    }
    if (best != (size_t)-1) {
        *p = c->cache[pick].first;
        c->cache[pick] = c->cache.back();
        c->cache.pop_back();
        c->cached -= best;
        c->live.emplace_back(*p, best);
        c->bytes += best;
        return PF_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && !c->cache.empty()) {
        (void)hipGetLastError();
        pf_cache_release(c);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return pf_set_hip_error(e, "hipMalloc", __FILE__, __LINE__);
    }
    c->live.emplace_back(*p, bytes);
    c->bytes += bytes;
    return PF_OK;
}

void pf_free(pf_ctx *c, void *p, size_t bytes)
{
    (void)bytes;
    if (!p) return;
    for (size_t k = 0; k < c->live.size(); ++k) {
        if (c->live[k].first == p) {
            const size_t sz = c->live[k].second;
            c->live[k] = c->live.back();
            c->live.pop_back();
            c->bytes -= std::min(c->bytes, sz);
            c->cache.emplace_back(p, sz);
            c->cached += sz;
            return;
        }
    }
    (void)hipFree(p);   // not from this context's allocator
}

void pf_cache_release(pf_ctx *c)
{
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &b : c->cache) (void)hipFree(b.first);
    c->cache.clear();
    c->cached = 0;
}

template <class T>
static void free_arr(pf_ctx *c, T *&p, size_t n)
{
    pf_free(c, p, n * sizeof(T));
    p = nullptr;
}

bool pf_timing_begin(pf_ctx *c, int k, hipEvent_t *start, hipEvent_t *stop)
{
    if (!c->timing) return false;
    while (c->ev_pool.size() < c->ev_next + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        c->ev_pool.push_back(e);
    }
    const int a = (int)c->ev_next, b = (int)c->ev_next + 1;
    c->ev_next += 2;
    c->ev_used[k].emplace_back(a, b);
    *start = c->ev_pool[a];
    *stop = c->ev_pool[b];
    return true;
}

int pf_grid_for(int64_t n)
{
    int64_t g = (n + PF_BLOCK - 1) / PF_BLOCK;
    if (g < 1) g = 1;
    if (g > PF_REDUCE_MAX_GRID) g = PF_REDUCE_MAX_GRID;
    return (int)g;
}

static const size_t kPartials = (size_t)PF_NQ * PF_REDUCE_MAX_GRID + 8;

void pf_graph_drop(pf_ctx *c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_iters = 0;
}

static void free_problem_vectors(pf_ctx *c)
{
    pf_graph_drop(c);
    pf_ranges_drop(c);
    pf_block_free(c);
    size_t nv = (size_t)c->neq + 1;
    free_arr(c, c->x, nv); free_arr(c, c->r, nv); free_arr(c, c->p, nv);
    free_arr(c, c->d, nv); free_arr(c, c->u, nv); free_arr(c, c->dinv, nv);
    free_arr(c, c->diag, nv); free_arr(c, c->diag0, nv); free_arr(c, c->f, nv);
    free_arr(c, c->fixstore, nv);
    free_arr(c, c->p32, nv);
    c->nfix = 0;
    c->pcg_ready = false;
}

static void free_matrices(pf_ctx *c)
{
    pf_graph_drop(c);
    pf_block_free(c);
    const size_t ns = (size_t)pf_sym_per(c->ntot) * c->nels;
    free_arr(c, c->K64, pf_k64_count(c));
    free_arr(c, c->K64b, (size_t)c->ntot * c->ntot * c->nels);
    free_arr(c, c->K32hi, ns);
    free_arr(c, c->K32lo, ns);
    c->have_K = false;
    c->k_shared = false;
    c->precon_given = false;
}

// Drop every buffer of the current problem (a context can be reused for a
// new mesh of another size).
static void reset_problem(pf_ctx *c)
{
    (void)hipStreamSynchronize(c->stream);
    free_problem_vectors(c);
    free_matrices(c);
    free_arr(c, c->g_coord, (size_t)c->nod * 3 * c->nels);
    free_arr(c, c->g_g, (size_t)c->ntot * c->nels);
    free_arr(c, c->csr_ptr, (size_t)c->neq + 1);
    free_arr(c, c->csr_slot, (size_t)c->nnz);
    free_arr(c, c->csr_slot_t, (size_t)c->nnz);
    free_arr(c, c->utemp, (size_t)c->ntot * c->nels);
    free_arr(c, c->utemp32, (size_t)c->ntot * c->nels);
    c->nod = c->nels = c->ntot = c->neq = 0;
    c->nnz = 0;
}

// A new steering of ntot x nels slots and neq equations: the previous
// problem's steering, maps and vectors freed (matrices too when ntot
// changes), g_g, csr_ptr, utemp and the PCG vectors allocated, vectors
// zeroed.  The caller fills g_g and builds the transposed steering.
int pf_steer_install(pf_ctx *c, int ntot, int nels, int neq)
{
    if (ntot != 60 && ntot != 24 && ntot != 8) return PF_ERR_ARG;   // hex20/hex8 elastic, hex8 Laplace
    if (c->nels && c->nels != nels) return PF_ERR_ARG;             // pf_mesh_set first for a new mesh
    const size_t nslot = (size_t)ntot * nels;
    if (nslot >= (size_t)INT32_MAX) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    PF_HIP(hipStreamSynchronize(c->stream));
    if (c->ntot && c->ntot != ntot) free_matrices(c);
    free_problem_vectors(c);
    free_arr(c, c->g_g, (size_t)c->ntot * c->nels);
    free_arr(c, c->csr_ptr, (size_t)c->neq + 1);
    free_arr(c, c->csr_slot, (size_t)c->nnz);
    free_arr(c, c->csr_slot_t, (size_t)c->nnz);
    free_arr(c, c->utemp, (size_t)c->ntot * c->nels);
    free_arr(c, c->utemp32, (size_t)c->ntot * c->nels);
    c->ntot = ntot;
    c->nels = nels;
    c->neq = neq;
    c->nnz = 0;
    int rc = pf_alloc(c, (void **)&c->g_g, nslot * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&c->csr_ptr, ((size_t)neq + 1) * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&c->utemp, nslot * sizeof(double));
    const size_t nv = (size_t)neq + 1;
    double **vecs[] = {&c->x, &c->r, &c->p, &c->d, &c->u, &c->dinv, &c->diag, &c->diag0, &c->f};
    for (double **v : vecs)
        if (!rc) rc = pf_alloc(c, (void **)v, nv * sizeof(double));
    if (!rc)
        for (double **v : vecs)
            if (hipMemsetAsync(*v, 0, nv * sizeof(double), c->stream) != hipSuccess) rc = PF_ERR_HIP;
    if (rc) pf_steer_drop(c);
    return rc;
}

// Forget the steering (after a failed set-up): maps, vectors and matrices
// freed, the mesh kept.
void pf_steer_drop(pf_ctx *c)
{
    (void)hipStreamSynchronize(c->stream);
    free_problem_vectors(c);
    free_matrices(c);
    free_arr(c, c->g_g, (size_t)c->ntot * c->nels);
    free_arr(c, c->csr_ptr, (size_t)c->neq + 1);
    free_arr(c, c->csr_slot, (size_t)c->nnz);
    free_arr(c, c->csr_slot_t, (size_t)c->nnz);
    free_arr(c, c->utemp, (size_t)c->ntot * c->nels);
    free_arr(c, c->utemp32, (size_t)c->ntot * c->nels);
    c->ntot = c->neq = 0;
    c->nnz = 0;
    if (!c->g_coord) c->nels = 0;
}

extern "C" {

int pf_version(void) { return 200; }

const char *pf_status_string(int s)
{
    switch (s) {
    case PF_OK: return "ok";
    case PF_ERR_ARG: return "bad argument or call order";
    case PF_ERR_HIP: return g_last_error[0] ? g_last_error : "HIP runtime error";
    case PF_ERR_OOM: return "device out of memory";
    case PF_ERR_STATE: return "missing set-up call";
    case PF_NOT_CONVERGED: return "PCG reached its iteration limit";
    case PF_ERR_BREAKDOWN: return "PCG breakdown (non-finite or zero curvature)";
    case PF_ERR_NODEVICE: return "no HIP device";
    case PF_ERR_TOO_LARGE: return "problem exceeds the allocation-free scratch (PF_NOALLOC_SCRATCH_BYTES)";
    default: return "unknown status";
    }
}

int pf_reduce_grid(int64_t n) { return pf_grid_for(n); }

int pf_create(int device, pf_ctx **out)
{
    if (!out) return PF_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return PF_ERR_NODEVICE;
    }
    if (device < 0 || device >= ndev) return PF_ERR_NODEVICE;
    PF_HIP(hipSetDevice(device));
    pf_ctx *c = new pf_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return pf_set_hip_error(e, "hipStreamCreate", __FILE__, __LINE__);
    }
    c->stream = c->own_stream;
    int rc = pf_scratch_init(device);   // device address of the allocation-free scratch
    if (!rc) rc = pf_alloc(c, (void **)&c->state, sizeof(pf_state));
    if (!rc) rc = pf_alloc(c, (void **)&c->partials, sizeof(double) * kPartials);
    if (!rc && hipHostMalloc((void **)&c->state_host, sizeof(pf_state), hipHostMallocDefault) != hipSuccess) rc = PF_ERR_OOM;
    if (!rc && hipHostMalloc((void **)&c->scalar_host, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) rc = PF_ERR_OOM;
    if (!rc && hipHostMalloc((void **)&c->state_query, sizeof(pf_state), hipHostMallocDefault) != hipSuccess) rc = PF_ERR_OOM;
    if (!rc && hipEventCreateWithFlags(&c->ev_query, hipEventDisableTiming) != hipSuccess) rc = PF_ERR_HIP;
    if (!rc) {
        hipError_t z = hipMemsetAsync(c->state, 0, sizeof(pf_state), c->stream);
        if (z == hipSuccess) z = hipStreamSynchronize(c->stream);
        if (z != hipSuccess) rc = pf_set_hip_error(z, "init state", __FILE__, __LINE__);
    }
    if (rc) {
        pf_destroy(c);
        return rc;
    }
    *out = c;
    return PF_OK;
}

int pf_destroy(pf_ctx *c)
{
    if (!c) return PF_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    reset_problem(c);
    free_arr(c, c->partials, kPartials);
    pf_free(c, c->state, sizeof(pf_state));
    if (c->state_host) (void)hipHostFree(c->state_host);
    if (c->scalar_host) (void)hipHostFree(c->scalar_host);
    if (c->state_query) (void)hipHostFree(c->state_query);
    if (c->ev_query) (void)hipEventDestroy(c->ev_query);
    pf_graph_drop(c);
    pf_cache_release(c);
    for (auto &b : c->live) (void)hipFree(b.first);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return PF_OK;
}

int pf_reset(pf_ctx *c)
{
    if (!c) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    reset_problem(c);
    c->query_pending = false;
    return PF_OK;
}

int pf_set_stream(pf_ctx *c, void *s)
{
    if (!c) return PF_ERR_ARG;
    PF_HIP(hipStreamSynchronize(c->stream));
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return PF_OK;
}

void *pf_get_stream(pf_ctx *c) { return c ? (void *)c->stream : nullptr; }

int pf_sync(pf_ctx *c)
{
    if (!c) return PF_ERR_ARG;
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_dev_alloc(pf_ctx *c, size_t bytes, void **dptr)
{
    if (!c || !dptr) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    return pf_alloc(c, dptr, bytes);
}

int pf_dev_free(pf_ctx *c, void *dptr)
{
    if (!c) return PF_ERR_ARG;
    if (!dptr) return PF_OK;
    PF_HIP(hipStreamSynchronize(c->stream));
    size_t sz = 0;
    if (hipMemPtrGetInfo(dptr, &sz) != hipSuccess) {
        (void)hipGetLastError();
        sz = 0;
    }
    pf_free(c, dptr, sz);
    return PF_OK;
}

int pf_memcpy_h2d(pf_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes)) return PF_ERR_ARG;
    PF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_memcpy_d2h(pf_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes)) return PF_ERR_ARG;
    PF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_mem_used(pf_ctx *c, size_t *bytes)
{
    if (!c || !bytes) return PF_ERR_ARG;
    *bytes = c->bytes + c->cached;
    return PF_OK;
}

int pf_tune_set(pf_ctx *c, const pf_tune *t)
{
    if (!c || !t) return PF_ERR_ARG;
    const int e = t->ebe_epl, w = t->blk_waves, r = t->blk_rows;
    if (e < 0 || e > 8 || (e & (e - 1))) return PF_ERR_ARG;
    if (c->ntot && e > pf_epl_max(c->ntot)) return PF_ERR_ARG;
    if (w < 0 || w > 4 || (w & (w - 1)) || r < 0 || r > 4 || (r & (r - 1))) return PF_ERR_ARG;
    pf_graph_drop(c);
    c->tune_epl = e;
    c->tune_blk_waves = w;
    c->tune_blk_rows = r;
    return PF_OK;
}

int pf_tune_get(pf_ctx *c, pf_tune *t)
{
    if (!c || !t) return PF_ERR_ARG;
    t->ebe_epl = c->tune_epl;
    t->blk_waves = c->tune_blk_waves;
    t->blk_rows = c->tune_blk_rows;
    return PF_OK;
}

int pf_device_info(pf_ctx *c, char *arch, int len, int *cus)
{
    if (!c || !arch || len < 1) return PF_ERR_ARG;
    hipDeviceProp_t pr;
    PF_HIP(hipGetDeviceProperties(&pr, c->device));
    snprintf(arch, (size_t)len, "%s", pr.gcnArchName);
    if (cus) *cus = pr.multiProcessorCount;
    return PF_OK;
}

int pf_trim(pf_ctx *c)
{
    if (!c) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    pf_cache_release(c);
    return PF_OK;
}

int pf_mesh_set(pf_ctx *c, int nod, int nels, const double *g_coord)
{
    if (!c || !g_coord || (nod != 8 && nod != 20) || nels <= 0) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    if (c->nels && (c->nels != nels || (c->ntot && c->ntot % nod != 0))) reset_problem(c);
    free_arr(c, c->g_coord, (size_t)c->nod * 3 * c->nels);
    c->nod = nod;
    c->nels = nels;
    size_t n = (size_t)nod * 3 * nels;
    int rc = pf_alloc(c, (void **)&c->g_coord, n * sizeof(double));
    if (rc) return rc;
    PF_HIP(hipMemcpyAsync(c->g_coord, g_coord, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_steer_set(pf_ctx *c, int ntot, int nels, int neq, const int *g_g)
{
    if (!c || !g_g || nels <= 0 || neq <= 0) return PF_ERR_ARG;
    int rc = pf_steer_install(c, ntot, nels, neq);
    if (rc) return rc;
    const size_t nslot = (size_t)ntot * nels;
    hipError_t he = hipMemcpyAsync(c->g_g, g_g, nslot * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (he != hipSuccess) {
        pf_steer_drop(c);
        return pf_set_hip_error(he, "steering upload", __FILE__, __LINE__);
    }
    rc = pf_csr_build(c);
    if (rc) pf_steer_drop(c);
    return rc;
}

static int alloc_k64(pf_ctx *c, int store, bool shared = false)
{
    if (store != PF_STORE_FP64 && store != PF_STORE_FP32) return PF_ERR_ARG;
    if (shared && store != PF_STORE_FP64) return PF_ERR_ARG;
    free_matrices(c);
    c->store = store;
    c->k_shared = shared;
    return pf_alloc(c, (void **)&c->K64, pf_k64_count(c) * sizeof(double));
}

// Storage for an assembly on the device: K64 for FP64, the packed hi/lo
// matrices (and the FP32 products' utemp32) for FP32 -- the FP32 assembly
// writes them directly, with no FP64 copy.
static int alloc_assembly(pf_ctx *c, int store)
{
    if (store == PF_STORE_FP64) return alloc_k64(c, store);
    if (store != PF_STORE_FP32) return PF_ERR_ARG;
    free_matrices(c);
    c->store = store;
    c->k_shared = false;
    const size_t ns = (size_t)pf_sym_per(c->ntot) * c->nels;
    int rc = pf_alloc(c, (void **)&c->K32hi, ns * sizeof(float));
    if (!rc) rc = pf_alloc(c, (void **)&c->K32lo, ns * sizeof(float));
    if (!rc && !c->utemp32) rc = pf_alloc(c, (void **)&c->utemp32, (size_t)c->ntot * c->nels * sizeof(float));
    return rc;
}

// After a direct FP32 assembly: the assembled diagonal from the element
// diagonals dg the assembly wrote (the same sums as from FP64 matrices).
static int finish_direct32(pf_ctx *c, const double *dg)
{
    const size_t nv = (size_t)c->neq + 1;
    int rc = pf_launch_diag_elem(c, dg, c->diag0);
    if (rc) return rc;
    PF_HIP(hipMemcpyAsync(c->diag, c->diag0, nv * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    c->have_K = true;
    c->pcg_ready = false;
    pf_ranges_drop(c);   // fixed rows are cleared
    free_arr(c, c->fixstore, nv);
    c->nfix = 0;
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

// After a FP64 assembly or upload into K64: the assembled diagonal
// (ascending element order, as SCATTER adds it) and, for FP32 storage, the
// hi/lo split, after which the FP64 copy is released.
static int finish_matrices(pf_ctx *c)
{
    const size_t nk = pf_k64_count(c);
    const size_t nv = (size_t)c->neq + 1;
    int rc = pf_launch_diag(c, c->diag0);
    if (rc) return rc;
    PF_HIP(hipMemcpyAsync(c->diag, c->diag0, nv * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (c->store == PF_STORE_FP32) {
        const size_t ns = (size_t)pf_sym_per(c->ntot) * c->nels;
        rc = pf_alloc(c, (void **)&c->K32hi, ns * sizeof(float));
        if (!rc) rc = pf_alloc(c, (void **)&c->K32lo, ns * sizeof(float));
        if (!rc && !c->utemp32) rc = pf_alloc(c, (void **)&c->utemp32, (size_t)c->ntot * c->nels * sizeof(float));
        if (!rc) rc = pf_launch_split_fp32(c);
        if (rc) return rc;
        PF_HIP(hipStreamSynchronize(c->stream));
        free_arr(c, c->K64, nk);
    }
    c->have_K = true;
    c->pcg_ready = false;
    pf_ranges_drop(c);   // fixed rows are cleared
    free_arr(c, c->fixstore, nv);
    c->nfix = 0;
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_assemble_elastic(pf_ctx *c, double e, double v, const double *e_elem, int store)
{
    if (!c) return PF_ERR_ARG;
    if (!c->g_coord || !c->g_g) return PF_ERR_STATE;
    if (c->ntot != 3 * c->nod) return PF_ERR_ARG;
    if (!(v > -1.0 && v < 0.5)) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    int rc = alloc_assembly(c, store);
    if (rc) return rc;
    double *ee = nullptr;
    if (e_elem) {
        rc = pf_alloc(c, (void **)&ee, sizeof(double) * c->nels);
        if (rc) return rc;
        PF_HIP(hipMemcpyAsync(ee, e_elem, sizeof(double) * c->nels, hipMemcpyHostToDevice, c->stream));
    }
    // DEEMAT (new_library.f90:1671-1686), evaluated with the same
    // expressions: dee = dee0 * e / (2(1+v)vv), dee0 = 1 | v2 | vv.
    const double v2 = v / (1.0 - v);
    const double vv = (1.0 - 2.0 * v) / (1.0 - v) * 0.5;
    const double den = 2.0 * (1.0 + v) * vv;
    // FP32 storage: the element diagonals go to utemp (ntot*nels doubles,
    // free until the PCG runs)
    rc = pf_launch_assemble_elastic(c, v2, vv, den, e, ee, c->utemp);
    if (ee) {
        (void)hipStreamSynchronize(c->stream);
        pf_free(c, ee, sizeof(double) * c->nels);
    }
    if (rc) return rc;
    return store == PF_STORE_FP32 ? finish_direct32(c, c->utemp) : finish_matrices(c);
}

int pf_assemble_laplace(pf_ctx *c, double kx, double ky, double kz, const double *k_elem, int store)
{
    if (!c) return PF_ERR_ARG;
    if (!c->g_coord || !c->g_g) return PF_ERR_STATE;
    if (c->ntot != c->nod || c->nod != 8) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    int rc = alloc_assembly(c, store);
    if (rc) return rc;
    double *kk = nullptr;
    if (k_elem) {
        rc = pf_alloc(c, (void **)&kk, sizeof(double) * c->nels);
        if (rc) return rc;
        PF_HIP(hipMemcpyAsync(kk, k_elem, sizeof(double) * c->nels, hipMemcpyHostToDevice, c->stream));
    }
    rc = pf_launch_assemble_laplace(c, kx, ky, kz, kk, c->utemp);
    if (kk) {
        (void)hipStreamSynchronize(c->stream);
        pf_free(c, kk, sizeof(double) * c->nels);
    }
    if (rc) return rc;
    return store == PF_STORE_FP32 ? finish_direct32(c, c->utemp) : finish_matrices(c);
}

int pf_assemble_transient(pf_ctx *c, double theta, double dtim, int np_types, const double *prop,
                          const int *etype, int store)
{
    if (!c || !prop || !etype || np_types < 1) return PF_ERR_ARG;
    if (!c->g_coord || !c->g_g) return PF_ERR_STATE;
    if (c->ntot != c->nod || c->nod != 8) return PF_ERR_ARG;
    for (int e = 0; e < c->nels; ++e)
        if (etype[e] < 1 || etype[e] > np_types) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    int rc = alloc_k64(c, store);
    if (!rc) rc = pf_alloc(c, (void **)&c->K64b, (size_t)64 * c->nels * sizeof(double));
    if (rc) return rc;
    double *dprop = nullptr;
    int *dtype = nullptr;
    rc = pf_alloc(c, (void **)&dprop, sizeof(double) * 5 * np_types);
    if (!rc) rc = pf_alloc(c, (void **)&dtype, sizeof(int) * c->nels);
    if (!rc) {
        hipError_t he = hipMemcpyAsync(dprop, prop, sizeof(double) * 5 * np_types, hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(dtype, etype, sizeof(int) * c->nels, hipMemcpyHostToDevice, c->stream);
        if (he != hipSuccess) rc = pf_set_hip_error(he, "transient upload", __FILE__, __LINE__);
    }
    if (!rc) rc = pf_launch_assemble_transient(c, theta, dtim, dprop, dtype);
    (void)hipStreamSynchronize(c->stream);
    pf_free(c, dprop, sizeof(double) * 5 * np_types);
    pf_free(c, dtype, sizeof(int) * c->nels);
    if (rc) return rc;
    return finish_matrices(c);
}

// storkm_pp is moved through a bounded staging buffer, chunk by chunk.
static long staging_elems(const pf_ctx *c)
{
    const long per = (long)c->ntot * c->ntot * (long)sizeof(double);
    long n = (256L << 20) / per;
    if (n < 1) n = 1;
    return n < c->nels ? n : c->nels;
}

int pf_storkm_set(pf_ctx *c, const double *storkm, int store)
{
    if (!c || !storkm) return PF_ERR_ARG;
    if (!c->g_g) return PF_ERR_STATE;
    PF_HIP(hipSetDevice(c->device));
    int rc = alloc_k64(c, store);
    if (rc) return rc;
    const long chunk = staging_elems(c), nn = (long)c->ntot * c->ntot;
    double *stage = nullptr;
    rc = pf_alloc(c, (void **)&stage, (size_t)chunk * nn * sizeof(double));
    for (long e0 = 0; !rc && e0 < c->nels; e0 += chunk) {
        const long nel = std::min(chunk, (long)c->nels - e0);
        hipError_t he = hipMemcpyAsync(stage, storkm + e0 * nn, (size_t)nel * nn * sizeof(double),
                                       hipMemcpyHostToDevice, c->stream);
        if (he != hipSuccess) rc = pf_set_hip_error(he, "storkm upload", __FILE__, __LINE__);
        if (!rc) rc = pf_launch_from_colmajor(c, stage, e0, nel);
    }
    (void)hipStreamSynchronize(c->stream);
    pf_free(c, stage, (size_t)chunk * nn * sizeof(double));
    if (rc) return rc;
    return finish_matrices(c);
}

int pf_storkm_shared(pf_ctx *c, const double *km)
{
    if (!c || !km) return PF_ERR_ARG;
    if (!c->g_g) return PF_ERR_STATE;
    PF_HIP(hipSetDevice(c->device));
    int rc = alloc_k64(c, PF_STORE_FP64, true);
    if (rc) return rc;
    const size_t nn = (size_t)c->ntot * c->ntot;
    double *stage = nullptr;
    rc = pf_alloc(c, (void **)&stage, nn * sizeof(double));
    if (!rc) {
        hipError_t he = hipMemcpyAsync(stage, km, nn * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (he != hipSuccess) rc = pf_set_hip_error(he, "km upload", __FILE__, __LINE__);
        if (!rc) rc = pf_launch_from_colmajor(c, stage, 0, 1);
    }
    (void)hipStreamSynchronize(c->stream);
    pf_free(c, stage, nn * sizeof(double));
    if (rc) return rc;
    return finish_matrices(c);
}

int pf_precon_set(pf_ctx *c, const double *diag_precon)
{
    if (!c || !diag_precon) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    PF_HIP(hipSetDevice(c->device));
    PF_HIP(hipMemcpyAsync(c->dinv + 1, diag_precon, sizeof(double) * c->neq, hipMemcpyHostToDevice, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    c->precon_given = true;
    c->pcg_ready = false;
    return PF_OK;
}

int pf_storkm_get(pf_ctx *c, double *storkm)
{
    if (!c || !storkm) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    PF_HIP(hipSetDevice(c->device));
    const long chunk = staging_elems(c), nn = (long)c->ntot * c->ntot;
    double *stage = nullptr;
    int rc = pf_alloc(c, (void **)&stage, (size_t)chunk * nn * sizeof(double));
    for (long e0 = 0; !rc && e0 < c->nels; e0 += chunk) {
        const long nel = std::min(chunk, (long)c->nels - e0);
        rc = pf_launch_to_colmajor(c, stage, e0, nel);
        if (rc) break;
        hipError_t he = hipMemcpyAsync(storkm + e0 * nn, stage, (size_t)nel * nn * sizeof(double),
                                       hipMemcpyDeviceToHost, c->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) rc = pf_set_hip_error(he, "storkm download", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(c->stream);
    pf_free(c, stage, (size_t)chunk * nn * sizeof(double));
    return rc;
}

int pf_diag_get(pf_ctx *c, double *diag)
{
    if (!c || !diag) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    PF_HIP(hipMemcpyAsync(diag, c->diag + 1, sizeof(double) * c->neq, hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_fixed_set(pf_ctx *c, int nfix, const int *no_f, double penalty, double *store_out)
{
    if (!c || nfix < 0 || (nfix && !no_f)) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    for (int i = 0; i < nfix; ++i)
        if (no_f[i] < 1 || no_f[i] > c->neq) return PF_ERR_ARG;
    pf_graph_drop(c);
    pf_ranges_drop(c);   // their fixed-row flags change
    PF_HIP(hipSetDevice(c->device));
    size_t nv = (size_t)c->neq + 1;
    free_arr(c, c->fixstore, nv);
    c->nfix = 0;
    c->pcg_ready = false;
    // Start from the element diagonal again so repeated calls do not stack.
    PF_HIP(hipMemcpyAsync(c->diag, c->diag0, nv * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (nfix == 0) {
        PF_HIP(hipStreamSynchronize(c->stream));
        return PF_OK;
    }
    std::vector<double> dg(nv);
    PF_HIP(hipMemcpyAsync(dg.data(), c->diag0, nv * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> fs(nv, 0.0);
    // p123.f90:120-125: diag(j) = diag(j) + penalty; store(i) = diag(j).
    for (int i = 0; i < nfix; ++i) {
        int j = no_f[i];
        dg[j] = dg[j] + penalty;
        fs[j] = dg[j];
        if (store_out) store_out[i] = dg[j];
    }
    int rc = pf_alloc(c, (void **)&c->fixstore, nv * sizeof(double));
    if (rc) return rc;
    PF_HIP(hipMemcpyAsync(c->diag, dg.data(), nv * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PF_HIP(hipMemcpyAsync(c->fixstore, fs.data(), nv * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    c->nfix = nfix;
    return PF_OK;
}

}  // extern "C"
