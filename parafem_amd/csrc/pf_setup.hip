// Problem set-up on the device: the element steering g_g_pp from the node
// freedom table, and the transposed steering (equation -> element slots)
// that the order-preserving scatter walks.
//
// Reference: REARRANGE + FIND_G3 / FIND_G4 (src/modules/shared/
// new_library.f90:3059-3271) give every free freedom of every node an
// equation number in ascending node, then freedom, order; g_g_pp(:,iel) is
// that table gathered through g_num_pp(:,iel) (p121.f90:44-47).  MAKE_GGL
// (src/modules/mpi/gather_scatter.f90:1387-1780) builds the halo maps; at one
// rank the only map left is the transposed steering, built here.
//
// The transposed steering lists, for equation q, the slots k = iel*ntot + i
// with g_g(i,iel) = q in ascending k.  Device construction: count the slots
// per equation (integer atomics: exact), exclusive scan of the counts, place
// every slot with an atomic cursor, then sort each equation's short list, so
// the result does not depend on the order the atomics ran in.
#include "pf_internal.h"

#include <algorithm>
#include <vector>

namespace {

constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_CHUNK = 256 * SCAN_ITEMS;

__global__ __launch_bounds__(256) void k_csr_count(const int *__restrict__ gg, long nslot, int neq,
                                                  int *__restrict__ cnt, int *__restrict__ err)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nslot; k += (long)gridDim.x * blockDim.x) {
        const int g = gg[k];
        if (g < 0 || g > neq) atomicOr(err, 1);
        else if (g) atomicAdd(&cnt[g - 1], 1);
    }
}

// Block sums of consecutive chunks of SCAN_CHUNK counts.
__global__ __launch_bounds__(256) void k_scan_sums(const int *__restrict__ cnt, int n, int *__restrict__ bsum)
{
    __shared__ int ws[4];
    const long base = (long)blockIdx.x * SCAN_CHUNK + (long)threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j)
        if (base + j < n) s += cnt[base + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Exclusive scan of the nb block sums in place (one block); total -> *total.
__global__ __launch_bounds__(1024) void k_scan_top(int *__restrict__ bsum, int nb, int *__restrict__ total)
{
    __shared__ int part[1024];
    const int per = (nb + 1023) / 1024;
    const int a = threadIdx.x * per, b = min(a + per, nb);
    int s = 0;
    for (int i = a; i < b; ++i) s += bsum[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < 1024; ++t) {
            const int v = part[t];
            part[t] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int i = a; i < b; ++i) {
        const int v = bsum[i];
        bsum[i] = run;
        run += v;
    }
}

// ptr[q + 1] = sum of cnt[0..q] (ptr[0] = 0 is set by the caller); flags an
// equation that no slot touches.
__global__ __launch_bounds__(256) void k_scan_apply(const int *__restrict__ cnt, int n, const int *__restrict__ boff,
                                                   int *__restrict__ ptr, int *__restrict__ err)
{
    __shared__ int ts[256];
    const long base = (long)blockIdx.x * SCAN_CHUNK + (long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        v[j] = base + j < n ? cnt[base + j] : 0;
        if (base + j < n && v[j] == 0) atomicOr(err, 2);
        s += v[j];
    }
    ts[threadIdx.x] = s;
    __syncthreads();
    // inclusive scan of the 256 thread sums (Hillis-Steele)
    for (int off = 1; off < 256; off <<= 1) {
        const int add = threadIdx.x >= off ? ts[threadIdx.x - off] : 0;
        __syncthreads();
        ts[threadIdx.x] += add;
        __syncthreads();
    }
    int run = boff[blockIdx.x] + ts[threadIdx.x] - s;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        run += v[j];
        if (base + j < n) ptr[base + j + 1] = run;
    }
}

__global__ __launch_bounds__(256) void k_csr_fill(const int *__restrict__ gg, long nslot, int *__restrict__ cursor,
                                                 int *__restrict__ slot)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nslot; k += (long)gridDim.x * blockDim.x) {
        const int g = gg[k];
        if (g) slot[atomicAdd(&cursor[g - 1], 1)] = (int)k;
    }
}

// Each equation's slots in ascending order (insertion sort of a short list).
__global__ __launch_bounds__(256) void k_csr_sort(const int *__restrict__ ptr, int *__restrict__ slot, int neq)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= neq) return;
    const int a = ptr[q], b = ptr[q + 1];
    for (int i = a + 1; i < b; ++i) {
        const int v = slot[i];
        int j = i - 1;
        while (j >= a && slot[j] > v) {
            slot[j + 1] = slot[j];
            --j;
        }
        slot[j + 1] = v;
    }
}

// slot_t[s] = i*nels + e for slot[s] = e*ntot + i: where the iteration's
// products of a hex8 mesh keep row i of element e (all elements' row i side
// by side, so the scatter's reads of one slot across consecutive equations
// fall on consecutive words).  The list order is unchanged.
__global__ __launch_bounds__(256) void k_csr_transpose(const int *__restrict__ slot, long nnz, int ntot, int nels,
                                                       int *__restrict__ slot_t)
{
    for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < nnz; s += (long)gridDim.x * blockDim.x) {
        const int v = slot[s], e = v / ntot, i = v - e * ntot;
        slot_t[s] = i * nels + e;
    }
}

// g_g(i, iel) = nf(d, g_num(n, iel)) for element node n, freedom d, i = n*nodof + d.
__global__ __launch_bounds__(256) void k_steer_gather(const int *__restrict__ g_num, const int *__restrict__ nf,
                                                     int *__restrict__ gg, long nslot, int nod, int nodof)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nslot; k += (long)gridDim.x * blockDim.x) {
        const long e = k / (nod * nodof);
        const int i = (int)(k - e * nod * nodof);
        const int node = g_num[e * nod + i / nodof];
        gg[k] = nf[(long)(node - 1) * nodof + i % nodof];
    }
}

int grid_for_stride(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 16384)); }

}  // namespace

int pf_scan_blocks(long n) { return (int)std::max<long>(1, (n + SCAN_CHUNK - 1) / SCAN_CHUNK); }

void pf_scan_launch(pf_ctx *c, const int *cnt, int n, int *bsum, int *ptr, int *total, int *zero_flag)
{
    const int nb = pf_scan_blocks(n);
    hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(256), 0, c->stream, cnt, n, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, c->stream, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, c->stream, cnt, n, bsum, ptr, zero_flag);
}

// Transposed steering of c->g_g (already on the device, c->neq and c->ntot,
// c->nels set) into c->csr_ptr / c->csr_slot; PF_ERR_ARG when an entry is
// outside 0..neq or an equation has no slot.
int pf_csr_build(pf_ctx *c)
{
    const long nslot = (long)c->ntot * c->nels;
    const int neq = c->neq;
    const int nb = (int)((neq + SCAN_CHUNK - 1) / SCAN_CHUNK);
    int *cnt = nullptr, *bsum = nullptr, *flags = nullptr;
    int rc = pf_alloc(c, (void **)&cnt, (size_t)neq * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&bsum, (size_t)std::max(nb, 1) * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&flags, 4 * sizeof(int));   // err, nnz
    auto release = [&]() {
        (void)hipStreamSynchronize(c->stream);
        pf_free(c, cnt, (size_t)neq * sizeof(int));
        pf_free(c, bsum, (size_t)std::max(nb, 1) * sizeof(int));
        pf_free(c, flags, 4 * sizeof(int));
    };
    if (rc) {
        release();
        return rc;
    }
    int host[2] = {0, 0};
    hipError_t he = hipMemsetAsync(cnt, 0, (size_t)neq * sizeof(int), c->stream);
    if (he == hipSuccess) he = hipMemsetAsync(flags, 0, 4 * sizeof(int), c->stream);
    if (he == hipSuccess) he = hipMemsetAsync(c->csr_ptr, 0, sizeof(int), c->stream);
    if (he == hipSuccess) {
        const int gs = grid_for_stride(nslot);
        hipLaunchKernelGGL(k_csr_count, dim3(gs), dim3(256), 0, c->stream, c->g_g, nslot, neq, cnt, flags);
        pf_scan_launch(c, cnt, neq, bsum, c->csr_ptr, flags + 1, flags);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipMemcpyAsync(host, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) {
        release();
        return pf_set_hip_error(he, "transposed steering", __FILE__, __LINE__);
    }
    if (host[0]) {
        release();
        return PF_ERR_ARG;
    }
    c->nnz = host[1];
    rc = pf_alloc(c, (void **)&c->csr_slot, (size_t)std::max<int64_t>(c->nnz, 1) * sizeof(int));
    if (!rc && c->ntot == 8)
        rc = pf_alloc(c, (void **)&c->csr_slot_t, (size_t)std::max<int64_t>(c->nnz, 1) * sizeof(int));
    if (rc) {
        release();
        return rc;
    }
    // the counts become the placement cursors: cursor[q] = ptr[q]
    he = hipMemcpyAsync(cnt, c->csr_ptr, (size_t)neq * sizeof(int), hipMemcpyDeviceToDevice, c->stream);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(k_csr_fill, dim3(grid_for_stride(nslot)), dim3(256), 0, c->stream, c->g_g, nslot, cnt,
                           c->csr_slot);
        hipLaunchKernelGGL(k_csr_sort, dim3((neq + 255) / 256), dim3(256), 0, c->stream, c->csr_ptr, c->csr_slot,
                           neq);
        if (c->csr_slot_t)
            hipLaunchKernelGGL(k_csr_transpose, dim3(grid_for_stride(c->nnz)), dim3(256), 0, c->stream, c->csr_slot,
                               (long)c->nnz, c->ntot, c->nels, c->csr_slot_t);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    release();
    if (he != hipSuccess) return pf_set_hip_error(he, "transposed steering", __FILE__, __LINE__);
    return PF_OK;
}

extern "C" int pf_steer_build(pf_ctx *c, int nod, int nels, int nodof, const int *g_num, int nn, int nr,
                              const int *rest, int *neq_out, int *nf_out)
{
    if (!c || !g_num || !neq_out || nod < 1 || nels < 1 || nodof < 1 || nn < 1 || nr < 0 || (nr && !rest))
        return PF_ERR_ARG;
    // node freedom table on the host: ascending node, then freedom; nodes
    // without a rest row are free; a later row for a node overrides an
    // earlier one
    std::vector<int> nf((size_t)nn * nodof, 1);
    for (int r = 0; r < nr; ++r) {
        const int *row = rest + (size_t)r * (nodof + 1);
        if (row[0] < 1 || row[0] > nn) return PF_ERR_ARG;
        for (int d = 0; d < nodof; ++d) nf[(size_t)(row[0] - 1) * nodof + d] = row[1 + d] != 0;
    }
    int64_t neq = 0;
    for (auto &v : nf) v = v ? (int)++neq : 0;
    if (neq < 1 || neq >= INT32_MAX) return PF_ERR_ARG;
    for (long k = 0; k < (long)nod * nels; ++k)
        if (g_num[k] < 1 || g_num[k] > nn) return PF_ERR_ARG;
    int rc = pf_steer_install(c, nod * nodof, nels, (int)neq);
    if (rc) return rc;
    // g_g gathered on the device, then the transposed steering
    const long nslot = (long)nod * nodof * nels;
    int *dnum = nullptr, *dnf = nullptr;
    rc = pf_alloc(c, (void **)&dnum, (size_t)nod * nels * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&dnf, nf.size() * sizeof(int));
    if (!rc) {
        hipError_t he = hipMemcpyAsync(dnum, g_num, (size_t)nod * nels * sizeof(int), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(dnf, nf.data(), nf.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess) {
            hipLaunchKernelGGL(k_steer_gather, dim3(grid_for_stride(nslot)), dim3(256), 0, c->stream, dnum, dnf,
                               c->g_g, nslot, nod, nodof);
            he = hipGetLastError();
        }
        if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) rc = pf_set_hip_error(he, "steering", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(c->stream);
    pf_free(c, dnum, (size_t)nod * nels * sizeof(int));
    pf_free(c, dnf, nf.size() * sizeof(int));
    if (!rc) rc = pf_csr_build(c);
    if (rc) {
        pf_steer_drop(c);
        return rc;
    }
    *neq_out = (int)neq;
    if (nf_out) std::copy(nf.begin(), nf.end(), nf_out);
    return PF_OK;
}

extern "C" int pf_steer_get(pf_ctx *c, int *g_g_pp)
{
    if (!c || !g_g_pp) return PF_ERR_ARG;
    if (!c->g_g) return PF_ERR_STATE;
    PF_HIP(hipMemcpyAsync(g_g_pp, c->g_g, (size_t)c->ntot * c->nels * sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}
