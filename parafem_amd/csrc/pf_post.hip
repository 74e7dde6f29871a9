// Post-solve stress recovery: the stresses at every element's centroid.
//
// Reference: src/programs/5th_ed/p121/p121.f90:112-123 (done there for
// element 1 only): eld = GATHER(xnew), points = 0 and nip = 1, SHAPE_DER,
// jac = MATMUL(der, coord), INVERT(jac), deriv = MATMUL(jac, der),
// BEEMAT(bee, deriv), eps = MATMUL(bee, eld), sigma = MATMUL(dee, eps).
// Element routines: src/modules/shared/new_library.f90 (SHAPE_DER :490-912,
// BEEMAT :918-1000, DEEMAT :1604-1691), src/modules/mpi/maths.f90 (INVERT
// :490-565).  One thread per element; every sum in the reference's index
// order (products with B's structural zeros skipped, which adds exact
// zeros), no contraction, so the result equals the sequential evaluation.
#include "pf_internal.h"

#include <vector>

namespace {

// SHAPE_DER at (0,0,0): derivative a of shape function l.
template <int NOD>
__device__ __forceinline__ double der0(int a, int l)
{
    if constexpr (NOD == 8) {
        // case(8) with xi = eta = zeta = 0: every factor is 1, signs below
        constexpr int s[3][8] = {{-1, -1, 1, 1, -1, -1, 1, 1},
                                 {-1, -1, -1, -1, 1, 1, 1, 1},
                                 {-1, 1, 1, -1, -1, 1, 1, -1}};
        return 0.125 * s[a][l];
    } else {
        constexpr int xii[20] = {-1, -1, -1, 0, 1, 1, 1, 0, -1, -1, 1, 1, -1, -1, -1, 0, 1, 1, 1, 0};
        constexpr int etai[20] = {-1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1};
        constexpr int zetai[20] = {-1, 0, 1, 1, 1, 0, -1, -1, -1, 1, 1, -1, -1, 0, 1, 1, 1, 0, -1, -1};
        const double sx = xii[l], sy = etai[l], sz = zetai[l];
        const double xi = 0.0, eta = 0.0, zeta = 0.0;
        const double xi0 = xi * sx, eta0 = eta * sy, zeta0 = zeta * sz;
        const int n = l + 1;
        double d[3];
        if (n == 4 || n == 8 || n == 16 || n == 20) {
            d[0] = -.5 * xi * (1. + eta0) * (1. + zeta0);
            d[1] = .25 * sy * (1. - xi * xi) * (1. + zeta0);
            d[2] = .25 * sz * (1. - xi * xi) * (1. + eta0);
        } else if (n >= 9 && n <= 12) {
            d[0] = .25 * sx * (1. - eta * eta) * (1. + zeta0);
            d[1] = -.5 * eta * (1. + xi0) * (1. + zeta0);
            d[2] = .25 * sz * (1. + xi0) * (1. - eta * eta);
        } else if (n == 2 || n == 6 || n == 14 || n == 18) {
            d[0] = .25 * sx * (1. + eta0) * (1. - zeta * zeta);
            d[1] = .25 * sy * (1. + xi0) * (1. - zeta * zeta);
            d[2] = -.5 * zeta * (1. + xi0) * (1. + eta0);
        } else {
            d[0] = .125 * sx * (1. + eta0) * (1. + zeta0) * (2. * xi0 + eta0 + zeta0 - 1.);
            d[1] = .125 * sy * (1. + xi0) * (1. + zeta0) * (xi0 + 2. * eta0 + zeta0 - 1.);
            d[2] = .125 * sz * (1. + xi0) * (1. + eta0) * (xi0 + eta0 + 2. * zeta0 - 1.);
        }
        return d[a];
    }
}

template <int NOD>
__global__ __launch_bounds__(128) void k_stress(const double *__restrict__ gc, const int *__restrict__ gg,
                                               const double *__restrict__ xext, double e, double v,
                                               const double *__restrict__ e_elem, double *__restrict__ sigma,
                                               int nels)
{
    const int iel = blockIdx.x * blockDim.x + threadIdx.x;
    if (iel >= nels) return;
    constexpr int NT = 3 * NOD;
    const double *coord = gc + (long)iel * NOD * 3;      // (nod,3) column-major
    // jac = MATMUL(der, coord)
    double m[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int l = 0; l < NOD; ++l) s = s + der0<NOD>(a, l) * coord[l + NOD * b];
            m[a][b] = s;
        }
    // INVERT (maths.f90:525-547)
    double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]);
    det = det - m[0][1] * (m[1][0] * m[2][2] - m[2][0] * m[1][2]);
    det = det + m[0][2] * (m[1][0] * m[2][1] - m[2][0] * m[1][1]);
    double ji[3][3];
    ji[0][0] = m[1][1] * m[2][2] - m[2][1] * m[1][2];
    ji[1][0] = -m[1][0] * m[2][2] + m[2][0] * m[1][2];
    ji[2][0] = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    ji[0][1] = -m[0][1] * m[2][2] + m[2][1] * m[0][2];
    ji[1][1] = m[0][0] * m[2][2] - m[2][0] * m[0][2];
    ji[2][1] = -m[0][0] * m[2][1] + m[2][0] * m[0][1];
    ji[0][2] = m[0][1] * m[1][2] - m[1][1] * m[0][2];
    ji[1][2] = -m[0][0] * m[1][2] + m[1][0] * m[0][2];
    ji[2][2] = m[0][0] * m[1][1] - m[1][0] * m[0][1];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) ji[a][b] = ji[a][b] / det;
    // eps = MATMUL(bee, eld) with bee from deriv = MATMUL(jac^-1, der):
    // node m's columns x = 3m-2, y = 3m-1, z = 3m carry (BEEMAT)
    //   row1: dx*ux  row2: dy*uy  row3: dz*uz
    //   row4: dy*ux, dx*uy   row5: dz*uy, dy*uz   row6: dz*ux, dx*uz
    double eps[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int l = 0; l < NOD; ++l) {
        double dv[3];
        for (int a = 0; a < 3; ++a) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s = s + ji[a][k] * der0<NOD>(k, l);
            dv[a] = s;
        }
        const int *g = gg + (long)iel * NT + 3 * l;
        const double ux = xext[g[0]], uy = xext[g[1]], uz = xext[g[2]];
        eps[0] = eps[0] + dv[0] * ux;
        eps[1] = eps[1] + dv[1] * uy;
        eps[2] = eps[2] + dv[2] * uz;
        eps[3] = eps[3] + dv[1] * ux;
        eps[3] = eps[3] + dv[0] * uy;
        eps[4] = eps[4] + dv[2] * uy;
        eps[4] = eps[4] + dv[1] * uz;
        eps[5] = eps[5] + dv[2] * ux;
        eps[5] = eps[5] + dv[0] * uz;
    }
    // sigma = MATMUL(dee, eps), DEEMAT(e, v)
    const double ee = e_elem ? e_elem[iel] : e;
    const double v2 = v / (1.0 - v);
    const double vv = (1.0 - 2.0 * v) / (1.0 - v) * 0.5;
    const double f = ee / (2.0 * (1.0 + v) * vv);
    double dee[6][6];
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) dee[a][b] = 0.0;
    for (int a = 0; a < 3; ++a) dee[a][a] = 1.0;
    for (int a = 3; a < 6; ++a) dee[a][a] = vv;
    dee[0][1] = dee[1][0] = dee[0][2] = dee[2][0] = dee[1][2] = dee[2][1] = v2;
    for (int a = 0; a < 6; ++a) {
        double s = 0.0;
        for (int t = 0; t < 6; ++t) s = s + dee[a][t] * f * eps[t];
        sigma[(long)iel * 6 + a] = s;
    }
}

// SCATTER_NODES at one rank: one thread per node adds the contributions of
// the node's (element, element-node) slots in ascending order, then divides
// by their number when averaging.
__global__ __launch_bounds__(256) void k_scatter_nodes(const int *__restrict__ ptr, const long *__restrict__ slot,
                                                      const double *__restrict__ contrib, double *__restrict__ out,
                                                      int nn, int numvar, int flag)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nn) return;
    const int a = ptr[node], b = ptr[node + 1];
    for (int k = 0; k < numvar; ++k) {
        double s = 0.0;
        for (int q = a; q < b; ++q) s = __dadd_rn(s, contrib[slot[q] * numvar + k]);
        if (flag == 1) s = s / (double)(b - a);
        out[(long)node * numvar + k] = s;
    }
}

}  // namespace

extern "C" int pf_scatter_nodes(pf_ctx *c, int nn, int nod, int nels, const int *g_num, int numvar,
                                const double *contrib, double *nodal, int flag)
{
    if (!c || nn < 1 || nod < 1 || nels < 0 || numvar < 1 || !nodal || (nels && (!g_num || !contrib)))
        return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    // node -> (element, element node) slots, ascending; nodes outside 1..nn
    // are not this rank's and are skipped
    const long ns = (long)nels * nod;
    std::vector<int> ptr((size_t)nn + 1, 0);
    for (long s = 0; s < ns; ++s)
        if (g_num[s] >= 1 && g_num[s] <= nn) ++ptr[g_num[s]];
    for (int i = 0; i < nn; ++i) ptr[i + 1] += ptr[i];
    std::vector<long> slot((size_t)std::max(ptr[nn], 1));
    {
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (long s = 0; s < ns; ++s)
            if (g_num[s] >= 1 && g_num[s] <= nn) slot[fill[g_num[s] - 1]++] = s;
    }
    const size_t nslot = slot.size(), nc = (size_t)std::max<long>(ns, 1) * numvar, no = (size_t)nn * numvar;
    int *dptr = nullptr;
    long *dslot = nullptr;
    double *dc = nullptr, *dout = nullptr;
    int rc = pf_alloc(c, (void **)&dptr, ((size_t)nn + 1) * sizeof(int));
    if (!rc) rc = pf_alloc(c, (void **)&dslot, nslot * sizeof(long));
    if (!rc) rc = pf_alloc(c, (void **)&dc, nc * sizeof(double));
    if (!rc) rc = pf_alloc(c, (void **)&dout, no * sizeof(double));
    if (!rc) {
        hipError_t he = hipMemcpyAsync(dptr, ptr.data(), ((size_t)nn + 1) * sizeof(int), hipMemcpyHostToDevice,
                                       c->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(dslot, slot.data(), nslot * sizeof(long), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess && ns)
            he = hipMemcpyAsync(dc, contrib, (size_t)ns * numvar * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess) {
            hipLaunchKernelGGL(k_scatter_nodes, dim3((nn + 255) / 256), dim3(256), 0, c->stream, dptr, dslot, dc,
                               dout, nn, numvar, flag);
            he = hipGetLastError();
        }
        if (he == hipSuccess) he = hipMemcpyAsync(nodal, dout, no * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) rc = pf_set_hip_error(he, "scatter_nodes", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(c->stream);
    pf_free(c, dptr, ((size_t)nn + 1) * sizeof(int));
    pf_free(c, dslot, nslot * sizeof(long));
    pf_free(c, dc, nc * sizeof(double));
    pf_free(c, dout, no * sizeof(double));
    return rc;
}

extern "C" int pf_stress_centroid(pf_ctx *c, const double *x, double e, double v, const double *e_elem,
                                  double *sigma)
{
    if (!c || !x || !sigma) return PF_ERR_ARG;
    if (!c->g_coord || !c->g_g) return PF_ERR_STATE;
    if (c->ntot != 3 * c->nod || (c->nod != 20 && c->nod != 8)) return PF_ERR_ARG;
    if (!(v > -1.0 && v < 0.5)) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    const size_t nv = (size_t)c->neq + 1, ns = (size_t)c->nels * 6;
    double *xe = nullptr, *sg = nullptr, *ee = nullptr;
    int rc = pf_alloc(c, (void **)&xe, nv * sizeof(double));
    if (!rc) rc = pf_alloc(c, (void **)&sg, ns * sizeof(double));
    if (!rc && e_elem) rc = pf_alloc(c, (void **)&ee, (size_t)c->nels * sizeof(double));
    if (!rc) {
        const double zero = 0.0;
        hipError_t he = hipMemcpyAsync(xe, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess) he = hipMemcpyAsync(xe + 1, x, c->neq * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess && e_elem)
            he = hipMemcpyAsync(ee, e_elem, c->nels * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (he != hipSuccess) rc = pf_set_hip_error(he, "stress upload", __FILE__, __LINE__);
    }
    if (!rc) {
        const int grid = (c->nels + 127) / 128;
        // the element matrices' e applies when no per-element modulus is given
        const double *eptr = e_elem ? ee : nullptr;
        (void)eptr;
        if (c->nod == 20)
            hipLaunchKernelGGL(k_stress<20>, dim3(grid), dim3(128), 0, c->stream, c->g_coord, c->g_g, xe, e, v, ee,
                               sg, c->nels);
        else
            hipLaunchKernelGGL(k_stress<8>, dim3(grid), dim3(128), 0, c->stream, c->g_coord, c->g_g, xe, e, v, ee,
                               sg, c->nels);
        hipError_t he = hipGetLastError();
        if (he == hipSuccess) he = hipMemcpyAsync(sigma, sg, ns * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) rc = pf_set_hip_error(he, "stress", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(c->stream);
    pf_free(c, xe, nv * sizeof(double));
    pf_free(c, sg, ns * sizeof(double));
    if (ee) pf_free(c, ee, (size_t)c->nels * sizeof(double));
    return rc;
}
