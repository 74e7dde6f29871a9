/*
 * CPU oracle for the EBE-PCG path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of ParaFEM's p121 (hex20 elasticity) and p123
 * (hex8 Laplace) element assembly and element-by-element PCG.  It is the
 * checker for the HIP library (tests/, __graft_entry__.smoke()) and the CPU
 * baseline leg of bench.py; it is never linked into or called by the product.
 *
 * Reference files (paths relative to the ParaFEM source tree):
 *   SAMPLE      src/modules/shared/new_library.f90:1006-1530 (hex nip=8 :1406-1433)
 *   SHAPE_DER   new_library.f90:490-912 (hex8 :769-794, hex20 :865-896)
 *   BEEMAT      new_library.f90:918-1000 (ih=6 :976-993)
 *   DEEMAT      new_library.f90:1604-1691 (3D :1671-1686)
 *   DETERMINANT src/modules/mpi/maths.f90:571-629, INVERT maths.f90:490-565
 *   p121 loops  src/programs/5th_ed/p121/p121.f90:54-104
 *   p123 loops  src/programs/5th_ed/p123/p123.f90:70-151
 *   GATHER/SCATTER  src/modules/mpi/gather_scatter.f90:547-850 (1 rank)
 *   DOT_PRODUCT_P maths.f90:168-216, SUM_P :271-315, checon_par :999-1069
 *
 * Arithmetic contract: compiled with -ffp-contract=off so every Fortran
 * expression is evaluated as written (no fused multiply-add), except the
 * element mat-vec, which accumulates u_i = fma(K_ij, p_j, u_i) for
 * j = 1..ntot in ascending order (MATMUL fixes no order; this is the order
 * the device kernel uses, so the two agree bit for bit on identical K, p).
 * The scatter adds element contributions in ascending element order, as
 * SCATTER does at one rank (gather_scatter.f90:758-761).  Dots are
 * sequential with nthreads == 1 (DOT_PRODUCT); with nthreads > 1 they are
 * per-thread static-chunk partial sums added in thread order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX2(i, j, ld) ((size_t)(i) + (size_t)(ld) * (size_t)(j))

/* ---------------------------------------------------------------- a1 */
/* SAMPLE('hexahedron', s, wt) with nip = 8: points (8,3) column-major. */
void or_sample_hex8(double *s, double *wt)
{
    const double r = 1.0 / sqrt(3.0);
    /* new_library.f90:1406-1433: points 6 and 7 are (-,-,+) and (-,+,-) */
    static const int sg[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
                                 {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    for (int i = 0; i < 8; ++i) {
        for (int k = 0; k < 3; ++k) s[IDX2(i, k, 8)] = sg[i][k] > 0 ? r : -r;
        wt[i] = 1.0;
    }
}

/* ---------------------------------------------------------------- a2 */
/* DEEMAT(dee,e,v) for nst = 6: dee (6,6) column-major. */
void or_deemat6(double *dee, double e, double v)
{
    memset(dee, 0, 36 * sizeof(double));
    double v2 = v / (1.0 - v);
    double vv = (1.0 - 2.0 * v) / (1.0 - v) * 0.5;
    for (int i = 0; i < 3; ++i) dee[IDX2(i, i, 6)] = 1.0;
    for (int i = 3; i < 6; ++i) dee[IDX2(i, i, 6)] = vv;
    dee[IDX2(0, 1, 6)] = v2; dee[IDX2(1, 0, 6)] = v2;
    dee[IDX2(0, 2, 6)] = v2; dee[IDX2(2, 0, 6)] = v2;
    dee[IDX2(1, 2, 6)] = v2; dee[IDX2(2, 1, 6)] = v2;
    double f = e / (2.0 * (1.0 + v) * vv);
    for (int k = 0; k < 36; ++k) dee[k] = dee[k] * f;
}

/* ---------------------------------------------------------------- a3 */
/* SHAPE_DER(der, points, i) for 3D nod = 8 or 20; der (3,nod) col-major,
 * points (nip,3) col-major, i is 1-based like the reference. */
int or_shape_der(double *der, const double *points, int nip, int i, int nod)
{
    double xi = points[IDX2(i - 1, 0, nip)];
    double eta = points[IDX2(i - 1, 1, nip)];
    double zeta = points[IDX2(i - 1, 2, nip)];
    if (nod == 8) {
        double etam = 1.0 - eta, xim = 1.0 - xi, zetam = 1.0 - zeta;
        double etap = eta + 1.0, xip = xi + 1.0, zetap = zeta + 1.0;
#define D(a, l) der[IDX2((a) - 1, (l) - 1, 3)]
        D(1, 1) = -0.125 * etam * zetam;  D(1, 2) = -0.125 * etam * zetap;
        D(1, 3) = 0.125 * etam * zetap;   D(1, 4) = 0.125 * etam * zetam;
        D(1, 5) = -0.125 * etap * zetam;  D(1, 6) = -0.125 * etap * zetap;
        D(1, 7) = 0.125 * etap * zetap;   D(1, 8) = 0.125 * etap * zetam;
        D(2, 1) = -0.125 * xim * zetam;   D(2, 2) = -0.125 * xim * zetap;
        D(2, 3) = -0.125 * xip * zetap;   D(2, 4) = -0.125 * xip * zetam;
        D(2, 5) = 0.125 * xim * zetam;    D(2, 6) = 0.125 * xim * zetap;
        D(2, 7) = 0.125 * xip * zetap;    D(2, 8) = 0.125 * xip * zetam;
        D(3, 1) = -0.125 * xim * etam;    D(3, 2) = 0.125 * xim * etam;
        D(3, 3) = 0.125 * xip * etam;     D(3, 4) = -0.125 * xip * etam;
        D(3, 5) = -0.125 * xim * etap;    D(3, 6) = 0.125 * xim * etap;
        D(3, 7) = 0.125 * xip * etap;     D(3, 8) = -0.125 * xip * etap;
#undef D
        return 0;
    }
    if (nod == 20) {
        static const int xii[20] = {-1, -1, -1, 0, 1, 1, 1, 0, -1, -1, 1, 1, -1, -1, -1, 0, 1, 1, 1, 0};
        static const int etai[20] = {-1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1};
        static const int zetai[20] = {-1, 0, 1, 1, 1, 0, -1, -1, -1, 1, 1, -1, -1, 0, 1, 1, 1, 0, -1, -1};
        for (int l = 1; l <= 20; ++l) {
            double xi0 = xi * xii[l - 1], eta0 = eta * etai[l - 1], zeta0 = zeta * zetai[l - 1];
            double *d = der + 3 * (l - 1);
            if (l == 4 || l == 8 || l == 16 || l == 20) {
                d[0] = -.5 * xi * (1. + eta0) * (1. + zeta0);
                d[1] = .25 * etai[l - 1] * (1. - xi * xi) * (1. + zeta0);
                d[2] = .25 * zetai[l - 1] * (1. - xi * xi) * (1. + eta0);
            } else if (l >= 9 && l <= 12) {
                d[0] = .25 * xii[l - 1] * (1. - eta * eta) * (1. + zeta0);
                d[1] = -.5 * eta * (1. + xi0) * (1. + zeta0);
                d[2] = .25 * zetai[l - 1] * (1. + xi0) * (1. - eta * eta);
            } else if (l == 2 || l == 6 || l == 14 || l == 18) {
                d[0] = .25 * xii[l - 1] * (1. + eta0) * (1. - zeta * zeta);
                d[1] = .25 * etai[l - 1] * (1. + xi0) * (1. - zeta * zeta);
                d[2] = -.5 * zeta * (1. + xi0) * (1. + eta0);
            } else {
                d[0] = .125 * xii[l - 1] * (1. + eta0) * (1. + zeta0) * (2. * xi0 + eta0 + zeta0 - 1.);
                d[1] = .125 * etai[l - 1] * (1. + xi0) * (1. + zeta0) * (xi0 + 2. * eta0 + zeta0 - 1.);
                d[2] = .125 * zetai[l - 1] * (1. + xi0) * (1. + eta0) * (xi0 + eta0 + 2. * zeta0 - 1.);
            }
        }
        return 0;
    }
    return 1;
}

/* ------------------------------------------------------------ a5, a6 */
double or_determinant3(const double *j)
{
#define J(a, b) j[IDX2((a) - 1, (b) - 1, 3)]
    double det = J(1, 1) * (J(2, 2) * J(3, 3) - J(3, 2) * J(2, 3));
    det = det - J(1, 2) * (J(2, 1) * J(3, 3) - J(3, 1) * J(2, 3));
    det = det + J(1, 3) * (J(2, 1) * J(3, 2) - J(3, 1) * J(2, 2));
    return det;
#undef J
}

void or_invert3(double *m)
{
#define M(a, b) m[IDX2((a) - 1, (b) - 1, 3)]
    double det = M(1, 1) * (M(2, 2) * M(3, 3) - M(3, 2) * M(2, 3));
    det = det - M(1, 2) * (M(2, 1) * M(3, 3) - M(3, 1) * M(2, 3));
    det = det + M(1, 3) * (M(2, 1) * M(3, 2) - M(3, 1) * M(2, 2));
    double j11 = M(2, 2) * M(3, 3) - M(3, 2) * M(2, 3);
    double j21 = -M(2, 1) * M(3, 3) + M(3, 1) * M(2, 3);
    double j31 = M(2, 1) * M(3, 2) - M(3, 1) * M(2, 2);
    double j12 = -M(1, 2) * M(3, 3) + M(3, 2) * M(1, 3);
    double j22 = M(1, 1) * M(3, 3) - M(3, 1) * M(1, 3);
    double j32 = -M(1, 1) * M(3, 2) + M(3, 1) * M(1, 2);
    double j13 = M(1, 2) * M(2, 3) - M(2, 2) * M(1, 3);
    double j23 = -M(1, 1) * M(2, 3) + M(2, 1) * M(1, 3);
    double j33 = M(1, 1) * M(2, 2) - M(2, 1) * M(1, 2);
    M(1, 1) = j11; M(1, 2) = j12; M(1, 3) = j13;
    M(2, 1) = j21; M(2, 2) = j22; M(2, 3) = j23;
    M(3, 1) = j31; M(3, 2) = j32; M(3, 3) = j33;
    for (int k = 0; k < 9; ++k) m[k] = m[k] / det;
#undef M
}

/* ---------------------------------------------------------------- a8 */
/* BEEMAT(bee, deriv) for ih = 6: bee (6, 3*nod) column-major. */
void or_beemat6(double *bee, const double *deriv, int nod)
{
    memset(bee, 0, sizeof(double) * 6 * 3 * nod);
    for (int m = 1; m <= nod; ++m) {
        int n = 3 * m, k = n - 1, l = k - 1;
        double x = deriv[IDX2(0, m - 1, 3)], y = deriv[IDX2(1, m - 1, 3)], z = deriv[IDX2(2, m - 1, 3)];
#define B(a, b) bee[IDX2((a) - 1, (b) - 1, 6)]
        B(1, l) = x; B(4, k) = x; B(6, n) = x;
        B(2, k) = y; B(4, l) = y; B(5, n) = y;
        B(3, n) = z; B(5, k) = z; B(6, l) = z;
#undef B
    }
}

/* jac = MATMUL(der, coord); det; invert; deriv = MATMUL(jac, der).
 * coord is g_coord_pp(:,:,iel): (nod,3) column-major. */
static double jac_deriv(const double *der, const double *coord, int nod, double *deriv)
{
    double jac[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int l = 0; l < nod; ++l) s = s + der[IDX2(a, l, 3)] * coord[IDX2(l, b, nod)];
            jac[IDX2(a, b, 3)] = s;
        }
    double det = or_determinant3(jac);
    or_invert3(jac);
    for (int a = 0; a < 3; ++a)
        for (int l = 0; l < nod; ++l) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s = s + jac[IDX2(a, k, 3)] * der[IDX2(k, l, 3)];
            deriv[IDX2(a, l, 3)] = s;
        }
    return det;
}

/* ---------------------------------------------------------------- a9 */
/* One hex20 (or hex8 elastic) element of p121.f90:56-64:
 * km += MATMUL(MATMUL(TRANSPOSE(bee),dee),bee)*det*weights(i). */
void or_elastic_km(const double *coord, int nod, const double *dee, double *km)
{
    int ntot = 3 * nod;
    double points[24], weights[8], der[60], deriv[60];
    double *bee = (double *)malloc(sizeof(double) * 6 * ntot);
    double *btd = (double *)malloc(sizeof(double) * ntot * 6);
    or_sample_hex8(points, weights);
    memset(km, 0, sizeof(double) * ntot * ntot);
    for (int gp = 1; gp <= 8; ++gp) {
        or_shape_der(der, points, 8, gp, nod);
        double det = jac_deriv(der, coord, nod, deriv);
        or_beemat6(bee, deriv, nod);
        for (int i = 0; i < ntot; ++i)
            for (int k = 0; k < 6; ++k) {
                double s = 0.0;
                for (int l = 0; l < 6; ++l) s = s + bee[IDX2(l, i, 6)] * dee[IDX2(l, k, 6)];
                btd[IDX2(i, k, ntot)] = s;
            }
        for (int j = 0; j < ntot; ++j)
            for (int i = 0; i < ntot; ++i) {
                double s = 0.0;
                for (int k = 0; k < 6; ++k) s = s + btd[IDX2(i, k, ntot)] * bee[IDX2(k, j, 6)];
                km[IDX2(i, j, ntot)] = km[IDX2(i, j, ntot)] + s * det * weights[gp - 1];
            }
    }
    free(bee);
    free(btd);
}

/* Centroid stresses, p121.f90:112-123 for every element: eld = GATHER(x),
 * SHAPE_DER at points = 0 (nip = 1), jac, INVERT, deriv, BEEMAT,
 * eps = MATMUL(bee, eld), sigma = MATMUL(dee, eps) with DEEMAT(e_iel, v).
 * sigma (6, nels) column-major. */
void or_centroid_stress(int nels, int nod, const double *g_coord, const int *g_g, const double *x,
                        double e, double v, const double *e_elem, double *sigma)
{
    int ntot = 3 * nod;
    double points[3] = {0.0, 0.0, 0.0}, der[60], deriv[60], eld[60], eps[6], dee[36];
    double *bee = (double *)malloc(sizeof(double) * 6 * ntot);
    or_shape_der(der, points, 1, 1, nod);
    for (int iel = 0; iel < nels; ++iel) {
        jac_deriv(der, g_coord + (size_t)iel * nod * 3, nod, deriv);
        or_beemat6(bee, deriv, nod);
        for (int j = 0; j < ntot; ++j) {
            int g = g_g[(size_t)iel * ntot + j];
            eld[j] = g ? x[g - 1] : 0.0;
        }
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
            for (int j = 0; j < ntot; ++j) s = s + bee[IDX2(r, j, 6)] * eld[j];
            eps[r] = s;
        }
        or_deemat6(dee, e_elem ? e_elem[iel] : e, v);
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
            for (int t = 0; t < 6; ++t) s = s + dee[IDX2(r, t, 6)] * eps[t];
            sigma[(size_t)iel * 6 + r] = s;
        }
    }
    free(bee);
}

/* SCATTER_NODES (gather_scatter.f90:1786-1929) at one rank: nodes_pp = nn,
 * node_start = 1, node_end = nn.  temp starts at zero; the loops run
 * iel = 1..nels, j = 1..nod, k = 1..numvar adding
 * element_contribution((j-1)*numvar+k, iel) into temp((g-1)*numvar+k) and
 * counting the terms; flag 1 then divides by the counts.  g_num (nod,nels)
 * column-major, 1-based; contrib (numvar*nod, nels); nodal (numvar*nn). */
void or_scatter_nodes(int nn, int nod, int nels, const int *g_num, int numvar, const double *contrib,
                      double *nodal, int flag)
{
    int *cnt = (int *)calloc((size_t)nn * numvar, sizeof(int));
    for (size_t i = 0; i < (size_t)nn * numvar; ++i) nodal[i] = 0.0;
    for (int iel = 0; iel < nels; ++iel)
        for (int j = 0; j < nod; ++j) {
            int g = g_num[(size_t)iel * nod + j];
            if (g < 1 || g > nn) continue;
            for (int k = 0; k < numvar; ++k) {
                size_t o = (size_t)(g - 1) * numvar + k;
                nodal[o] = nodal[o] + contrib[((size_t)iel * nod + j) * numvar + k];
                cnt[o] += 1;
            }
        }
    if (flag == 1)
        for (size_t i = 0; i < (size_t)nn * numvar; ++i) nodal[i] = nodal[i] / (double)cnt[i];
    free(cnt);
}

/* One hex8 Laplace element of p123.f90:70-84:
 * kc{x,y,z} += MATMUL(col,row)*det*w; storkc = kcx*kx + kcy*ky + kcz*kz. */
void or_laplace_kc(const double *coord, int nod, double kx, double ky, double kz, double *kc)
{
    double points[24], weights[8], der[60], deriv[60];
    double *kcx = (double *)calloc((size_t)nod * nod * 3, sizeof(double));
    double *kcy = kcx + nod * nod, *kcz = kcy + nod * nod;
    or_sample_hex8(points, weights);
    for (int gp = 1; gp <= 8; ++gp) {
        or_shape_der(der, points, 8, gp, nod);
        double det = jac_deriv(der, coord, nod, deriv);
        double *kk[3] = {kcx, kcy, kcz};
        for (int a = 0; a < 3; ++a)
            for (int j = 0; j < nod; ++j)
                for (int i = 0; i < nod; ++i)
                    kk[a][IDX2(i, j, nod)] = kk[a][IDX2(i, j, nod)] +
                        deriv[IDX2(a, i, 3)] * deriv[IDX2(a, j, 3)] * det * weights[gp - 1];
    }
    for (int k = 0; k < nod * nod; ++k) kc[k] = kcx[k] * kx + kcy[k] * ky + kcz[k] * kz;
    free(kcx);
}

/* SHAPE_FUN(fun, points, i) for the 8-node brick (new_library.f90:396-420):
 * fun = 0.125 * (1 -/+ xi) (1 -/+ eta) (1 -/+ zeta) per node, evaluated left
 * to right as written there.  points (nip,3) column-major, i 1-based. */
void or_shape_fun_hex8(double *fun, const double *points, int nip, int i)
{
    const double xi = points[IDX2(i - 1, 0, nip)], eta = points[IDX2(i - 1, 1, nip)];
    const double zeta = points[IDX2(i - 1, 2, nip)];
    const double etam = 1.0 - eta, xim = 1.0 - xi, zetam = 1.0 - zeta;
    const double etap = eta + 1.0, xip = xi + 1.0, zetap = zeta + 1.0;
    fun[0] = 0.125 * xim * etam * zetam;
    fun[1] = 0.125 * xim * etam * zetap;
    fun[2] = 0.125 * xip * etam * zetap;
    fun[3] = 0.125 * xip * etam * zetam;
    fun[4] = 0.125 * xim * etap * zetam;
    fun[5] = 0.125 * xim * etap * zetap;
    fun[6] = 0.125 * xip * etap * zetap;
    fun[7] = 0.125 * xip * etap * zetam;
}

/* One element of p124's loop (p124.f90:73-86), 8-node brick, nodof = 1:
 * per Gauss point kc = kc + MATMUL(MATMUL(TRANSPOSE(deriv),kay),deriv)*det
 * *weights(i) and pm = pm + MATMUL(TRANSPOSE(funny),funny)*det*weights(i)
 * *rho*cp; then ka = pm + kc*theta*dtim, kb = pm - kc*(1-theta)*dtim.
 * kay = diag(kay3); every MATMUL sums its terms in index order. */
void or_transient_km(const double *coord, const double *kay3, double rho, double cp, double theta, double dtim,
                     double *ka, double *kb)
{
    double points[24], weights[8], der[24], deriv[24], fun[8], kay[9], t1[24], kc[64], pm[64];
    for (int k = 0; k < 9; ++k) kay[k] = 0.0;
    for (int k = 0; k < 3; ++k) kay[IDX2(k, k, 3)] = kay3[k];
    for (int k = 0; k < 64; ++k) kc[k] = pm[k] = 0.0;
    or_sample_hex8(points, weights);
    for (int gp = 1; gp <= 8; ++gp) {
        or_shape_der(der, points, 8, gp, 8);
        or_shape_fun_hex8(fun, points, 8, gp);
        double det = jac_deriv(der, coord, 8, deriv);
        for (int a = 0; a < 8; ++a)            /* t1 = MATMUL(TRANSPOSE(deriv), kay): (8,3) */
            for (int c = 0; c < 3; ++c) {
                double s = 0.0;
                for (int m = 0; m < 3; ++m) s = s + deriv[IDX2(m, a, 3)] * kay[IDX2(m, c, 3)];
                t1[IDX2(a, c, 8)] = s;
            }
        for (int b = 0; b < 8; ++b)
            for (int a = 0; a < 8; ++a) {
                double s = 0.0;                /* MATMUL(t1, deriv) */
                for (int c = 0; c < 3; ++c) s = s + t1[IDX2(a, c, 8)] * deriv[IDX2(c, b, 3)];
                kc[IDX2(a, b, 8)] = kc[IDX2(a, b, 8)] + s * det * weights[gp - 1];
                double f = 0.0;                /* MATMUL(TRANSPOSE(funny), funny) */
                f = f + fun[a] * fun[b];
                pm[IDX2(a, b, 8)] = pm[IDX2(a, b, 8)] + f * det * weights[gp - 1] * rho * cp;
            }
    }
    for (int k = 0; k < 64; ++k) {
        ka[k] = pm[k] + kc[k] * theta * dtim;
        kb[k] = pm[k] - kc[k] * (1.0 - theta) * dtim;
    }
}

/* INVAR for nst = 6 (new_library.f90:1813-1900). */
static void or_invar6(const double *s, double *sigm, double *dsbar, double *theta)
{
    double sq3 = sqrt(3.0);
    double s1 = s[0], s2 = s[1], s3 = s[2], s4 = s[3], s5 = s[4], s6 = s[5];
    *sigm = (s1 + s2 + s3) / 3.0;
    double d2 = ((s1 - s2) * (s1 - s2) + (s2 - s3) * (s2 - s3) + (s3 - s1) * (s3 - s1)) / 6.0 + s4 * s4 + s5 * s5 +
                s6 * s6;
    double ds1 = s1 - *sigm, ds2 = s2 - *sigm, ds3 = s3 - *sigm;
    double d3 = ds1 * ds2 * ds3 - ds1 * s5 * s5 - ds2 * s6 * s6 - ds3 * s4 * s4 + 2.0 * s4 * s5 * s6;
    *dsbar = sq3 * sqrt(d2);
    if (*dsbar < 1.e-10) {
        *theta = 0.0;
    } else {
        double r = sqrt(d2);
        double sine = -(3.0 * sq3 * d3 / (2.0 * (r * r * r)));
        if (sine > 1.0) sine = 1.0;
        if (sine < -1.0) sine = -1.0;
        *theta = asin(sine) / 3.0;
    }
}

/* MOCOUF (new_library.f90:2364-2382): the Mohr-Coulomb yield function. */
double or_mocouf(double phi, double c, double sigm, double dsbar, double theta)
{
    double phir = phi * 4.0 * atan(1.0) / 180.0;
    double snph = sin(phir), csph = cos(phir), csth = cos(theta), snth = sin(theta);
    return snph * sigm + dsbar * (csth / sqrt(3.0) - snth * snph / 3.0) - c * csph;
}

/* MOCOUQ (new_library.f90:2423-2460): derivatives of the plastic potential. */
void or_mocouq(double psi, double dsbar, double theta, double *dq)
{
    double psir = psi * 4.0 * atan(1.0) / 180.0;
    double snth = sin(theta), snps = sin(psir), sq3 = sqrt(3.0);
    dq[0] = snps;
    if (fabs(snth) > 0.49) {
        double c1 = 1.0;
        if (snth < 0.0) c1 = -1.0;
        dq[1] = (sq3 * 0.5 - c1 * snps * 0.5 / sq3) * sq3 * 0.5 / dsbar;
        dq[2] = 0.0;
    } else {
        double csth = cos(theta), cs3th = cos(3.0 * theta), tn3th = tan(3.0 * theta);
        double tnth = snth / csth;
        dq[1] = sq3 * csth / dsbar * ((1.0 + tnth * tn3th) + snps * (tn3th - tnth) / sq3) * 0.5;
        dq[2] = 0.5 * 3.0 * (sq3 * snth + snps * csth) / (cs3th * dsbar * dsbar);
    }
}

/* FORMM for nst = 6 (new_library.f90:85-198); m (3,6,6): m1, m2, m3 with
 * entry (i,j) at i + 6 j. */
void or_formm6(const double *s, double *m)
{
    double *m1 = m, *m2 = m + 36, *m3 = m + 72;
    double sx = s[0], sy = s[1], sz = s[2], txy = s[3], tyz = s[4], tzx = s[5];
    double sigm = (sx + sy + sz) / 3.0;
    double dx = sx - sigm, dy = sy - sigm, dz = sz - sigm;
#define A(mm, i, j) mm[((i) - 1) + 6 * ((j) - 1)]
    for (int k = 0; k < 108; ++k) m[k] = 0.0;
    for (int i = 1; i <= 3; ++i)
        for (int j = 1; j <= 3; ++j) A(m1, i, j) = 1.0 / (3.0 * sigm);
    for (int i = 1; i <= 3; ++i) {
        A(m2, i, i) = 2.0;
        A(m2, i + 3, i + 3) = 6.0;
    }
    A(m2, 1, 2) = -1.0; A(m2, 1, 3) = -1.0; A(m2, 2, 3) = -1.0;
    A(m3, 1, 1) = dx; A(m3, 1, 2) = dz; A(m3, 1, 3) = dy; A(m3, 1, 4) = txy; A(m3, 1, 5) = -2.0 * tyz;
    A(m3, 1, 6) = tzx;
    A(m3, 2, 2) = dy; A(m3, 2, 3) = dx; A(m3, 2, 4) = txy; A(m3, 2, 5) = tyz; A(m3, 2, 6) = -2.0 * tzx;
    A(m3, 3, 3) = dz; A(m3, 3, 4) = -2.0 * txy; A(m3, 3, 5) = tyz; A(m3, 3, 6) = tzx;
    A(m3, 4, 4) = -3.0 * dz; A(m3, 4, 5) = 3.0 * tzx; A(m3, 4, 6) = 3.0 * tyz;
    A(m3, 5, 5) = -3.0 * dx; A(m3, 5, 6) = 3.0 * txy;
    A(m3, 6, 6) = -3.0 * dy;
    for (int i = 1; i <= 6; ++i)
        for (int j = i + 1; j <= 6; ++j) {
            A(m1, j, i) = A(m1, i, j);
            A(m2, j, i) = A(m2, i, j);
            A(m3, j, i) = A(m3, i, j);
        }
    for (int k = 0; k < 108; ++k) m[k] = m[k] / 3.0;
#undef A
}

/* p122's stress update over all elements (p122.f90:164-199), 8-node bricks:
 * x (neq) the displacement increment, evpt and tensor (6,8,nels), bload
 * (24,nels) out; final = plastic_converged .OR. plasiters == plasits. */
void or_plastic_update(int nels, const double *g_coord, const int *g_g, const double *x, double *evpt,
                       double *tensor, double *bload, double e, double v, double phi, double c, double psi,
                       double dt, int final_pass)
{
    double points[24], weights[8], der[24], deriv[24], dee[36], bee[144];
    or_sample_hex8(points, weights);
    or_deemat6(dee, e, v);
    for (int iel = 0; iel < nels; ++iel) {
        double eld[24], bl[24];
        for (int i = 0; i < 24; ++i) {
            int g = g_g[(size_t)iel * 24 + i];
            eld[i] = g ? x[g - 1] : 0.0;
            bl[i] = 0.0;
        }
        for (int ip = 1; ip <= 8; ++ip) {
            or_shape_der(der, points, 8, ip, 8);
            double det = jac_deriv(der, g_coord + (size_t)iel * 24, 8, deriv);
            or_beemat6(bee, deriv, 8);
            double eps[6], sigma[6], stress[6], devp[6] = {0, 0, 0, 0, 0, 0};
            double *ev = evpt + ((size_t)iel * 8 + ip - 1) * 6, *te = tensor + ((size_t)iel * 8 + ip - 1) * 6;
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int k = 0; k < 24; ++k) s = s + bee[IDX2(r, k, 6)] * eld[k];
                eps[r] = s;
            }
            for (int r = 0; r < 6; ++r) eps[r] = eps[r] - ev[r];
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int t = 0; t < 6; ++t) s = s + dee[IDX2(r, t, 6)] * eps[t];
                sigma[r] = s;
            }
            for (int r = 0; r < 6; ++r) stress[r] = sigma[r] + te[r];
            double sigm, dsbar, theta;
            or_invar6(stress, &sigm, &dsbar, &theta);
            double f = or_mocouf(phi, c, sigm, dsbar, theta);
            if (final_pass) {
                for (int r = 0; r < 6; ++r) devp[r] = stress[r];
            } else if (f >= 0.0) {
                double dq[3], m[108], evp[6];
                or_mocouq(psi, dsbar, theta, dq);
                or_formm6(stress, m);
                for (int a = 0; a < 6; ++a) {
                    double s = 0.0;
                    for (int b = 0; b < 6; ++b) {
                        double flow = f * (m[a + 6 * b] * dq[0] + m[36 + a + 6 * b] * dq[1] + m[72 + a + 6 * b] * dq[2]);
                        s = s + flow * stress[b];
                    }
                    evp[a] = s * dt;
                }
                for (int a = 0; a < 6; ++a) ev[a] = ev[a] + evp[a];
                for (int a = 0; a < 6; ++a) {
                    double s = 0.0;
                    for (int t = 0; t < 6; ++t) s = s + dee[IDX2(a, t, 6)] * evp[t];
                    devp[a] = s;
                }
            }
            if (f >= 0.0)
                for (int k = 0; k < 24; ++k) {
                    double s = 0.0;
                    for (int r = 0; r < 6; ++r) s = s + bee[IDX2(r, k, 6)] * devp[r];
                    bl[k] = bl[k] + s * det * weights[ip - 1];
                }
            if (final_pass)
                for (int r = 0; r < 6; ++r) te[r] = stress[r];
        }
        for (int i = 0; i < 24; ++i) bload[(size_t)iel * 24 + i] = bl[i];
    }
}

static int omp_threads(int nthreads)
{
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* p121 loop A over all elements.  g_coord is g_coord_pp (nod,3,nels);
 * e_elem (nullable) gives a per-element Young's modulus, else e. */
void or_assemble_elastic(int nels, int nod, const double *g_coord, const double *e_elem,
                         double e, double v, double *storkm, int nthreads)
{
    int ntot = 3 * nod;
    double dee0[36];
    or_deemat6(dee0, e, v);
#pragma omp parallel for schedule(static) num_threads(omp_threads(nthreads))
    for (long iel = 0; iel < nels; ++iel) {
        double dee[36];
        if (e_elem) or_deemat6(dee, e_elem[iel], v);
        else memcpy(dee, dee0, sizeof dee);
        or_elastic_km(g_coord + (size_t)iel * nod * 3, nod, dee, storkm + (size_t)iel * ntot * ntot);
    }
}

/* p124's loop over all elements: prop (5, np_types) = kx, ky, kz, rho, cp,
 * etype 1-based; storka and storkb (8,8,nels) column-major. */
void or_assemble_transient(int nels, const double *g_coord, const double *prop, const int *etype, double theta,
                           double dtim, double *storka, double *storkb, int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(omp_threads(nthreads))
    for (long iel = 0; iel < nels; ++iel) {
        const double *p = prop + 5 * (size_t)(etype[iel] - 1);
        or_transient_km(g_coord + (size_t)iel * 24, p, p[3], p[4], theta, dtim, storka + (size_t)iel * 64,
                        storkb + (size_t)iel * 64);
    }
}

/* p123 loop over all elements; k_elem (nullable) scales kx, ky, kz per element. */
void or_assemble_laplace(int nels, int nod, const double *g_coord, const double *k_elem,
                         double kx, double ky, double kz, double *storkc, int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(omp_threads(nthreads))
    for (long iel = 0; iel < nels; ++iel) {
        double s = k_elem ? k_elem[iel] : 1.0;
        or_laplace_kc(g_coord + (size_t)iel * nod * 3, nod, kx * s, ky * s, kz * s,
                      storkc + (size_t)iel * nod * nod);
    }
}

/* --------------------------------------------------------- a16 - a22 */
/* GATHER at one rank: pmul(:,e) = p(g_g(:,e)) with p(0) = 0. */
void or_gather(int nels, int ntot, const int *g_g, const double *p, double *pmul)
{
    for (size_t k = 0; k < (size_t)nels * ntot; ++k) pmul[k] = g_g[k] > 0 ? p[g_g[k] - 1] : 0.0;
}

/* SCATTER at one rank: u(g_g(:,e)) += utemp(:,e), ascending e. */
void or_scatter(int nels, int ntot, const int *g_g, const double *utemp, double *u)
{
    for (size_t k = 0; k < (size_t)nels * ntot; ++k)
        if (g_g[k] > 0) u[g_g[k] - 1] = u[g_g[k] - 1] + utemp[k];
}

/* Element products utemp(:,e) = storkm(:,:,e) . pmul(:,e). */
void or_elem_matvec(int nels, int ntot, const double *storkm, const double *pmul, double *utemp, int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(omp_threads(nthreads))
    for (long e = 0; e < nels; ++e) {
        const double *K = storkm + (size_t)e * ntot * ntot;
        const double *p = pmul + (size_t)e * ntot;
        double *u = utemp + (size_t)e * ntot;
        for (int i = 0; i < ntot; ++i) {
            double s = 0.0;
            for (int j = 0; j < ntot; ++j) s = fma(K[IDX2(i, j, ntot)], p[j], s);
            u[i] = s;
        }
    }
}

double or_dot(long n, const double *a, const double *b, int nthreads)
{
    int nt = omp_threads(nthreads);
    if (nt <= 1) {
        double s = 0.0;
        for (long i = 0; i < n; ++i) s = s + a[i] * b[i];
        return s;
    }
    double part[256];
    if (nt > 256) nt = 256;
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        long lo = n * t / nt, hi = n * (t + 1) / nt;
        double s = 0.0;
        for (long i = lo; i < hi; ++i) s = s + a[i] * b[i];
        part[t] = s;
    }
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s = s + part[t];
    return s;
}

double or_sum(long n, const double *a)
{
    double s = 0.0;
    for (long i = 0; i < n; ++i) s = s + a[i];
    return s;
}

/* checon_par: converged = max|xnew-x| / max|xnew| <= tol; x = xnew. */
int or_checon(long n, const double *xnew, double tol, double *x)
{
    double ml = 0.0, md = 0.0;
    for (long i = 0; i < n; ++i) {
        double a = fabs(xnew[i]), d = fabs(xnew[i] - x[i]);
        if (a > ml) ml = a;
        if (d > md) md = d;
        x[i] = xnew[i];
    }
    return (md / ml) <= tol;
}

/* Transposed steering for an order-preserving parallel scatter: for every
 * equation the (element, local dof) slots that touch it, ascending. */
typedef struct {
    long *ptr;
    long *slot;
} csr_t;

static csr_t build_csr(int nels, int ntot, int neq, const int *g_g)
{
    csr_t c;
    size_t nslot = (size_t)nels * ntot;
    c.ptr = (long *)calloc((size_t)neq + 1, sizeof(long));
    for (size_t k = 0; k < nslot; ++k)
        if (g_g[k] > 0) c.ptr[g_g[k]]++;
    for (long q = 0; q < neq; ++q) c.ptr[q + 1] += c.ptr[q];
    c.slot = (long *)malloc(sizeof(long) * (c.ptr[neq] ? c.ptr[neq] : 1));
    long *fill = (long *)malloc(sizeof(long) * ((size_t)neq + 1));
    memcpy(fill, c.ptr, sizeof(long) * ((size_t)neq + 1));
    for (size_t k = 0; k < nslot; ++k)
        if (g_g[k] > 0) c.slot[fill[g_g[k] - 1]++] = (long)k;
    free(fill);
    return c;
}

enum { OR_STOP_CHECON = 0, OR_STOP_RELRES = 1 };
enum { OR_DOT_SEQ = 0, OR_DOT_THREADS = 1, OR_DOT_DEVICE = 2 };

/* Summation order of the device library's FP64 reductions, restated on the
 * CPU so that a device solve can be compared bit for bit (it is not the
 * reference's order; the reference sums sequentially at one rank and by
 * contiguous per-rank chunks at several).  G = clamp(ceil(n/256), 1, 16384)
 * blocks of 256 threads; thread g sums terms g, g+256G, ...; each 64-lane
 * group folds with offsets 32..1; a block adds its 4 group sums left to
 * right; the G block sums are folded the same way by one block. */
static int dev_grid(long n)
{
    long g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > 16384) g = 16384;
    return (int)g;
}

static double dev_fold(double *v)
{
    for (int off = 32; off > 0; off >>= 1)
        for (int l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
    return v[0];
}

double or_dot_dev(long n, const double *a, const double *b)
{
    int G = dev_grid(n);
    long stride = (long)G * 256;
    double *P = (double *)malloc(sizeof(double) * G);
    double v[64], w[4];
    for (int blk = 0; blk < G; ++blk) {
        for (int wv = 0; wv < 4; ++wv) {
            for (int l = 0; l < 64; ++l) {
                double s = 0.0;
                for (long i = (long)blk * 256 + wv * 64 + l; i < n; i += stride)
                    s = s + (b ? a[i] * b[i] : a[i]);
                v[l] = s;
            }
            w[wv] = dev_fold(v);
        }
        P[blk] = ((w[0] + w[1]) + w[2]) + w[3];
    }
    for (int wv = 0; wv < 4; ++wv) {
        for (int l = 0; l < 64; ++l) {
            double s = 0.0;
            for (int k = wv * 64 + l; k < G; k += 256) s = s + P[k];
            v[l] = s;
        }
        w[wv] = dev_fold(v);
    }
    free(P);
    return ((w[0] + w[1]) + w[2]) + w[3];
}

static double dot_mode(long n, const double *a, const double *b, int nt, int order)
{
    if (order == OR_DOT_DEVICE) return or_dot_dev(n, a, b);
    return or_dot(n, a, b, order == OR_DOT_SEQ ? 1 : nt);
}

/* The p121/p123 PCG loop (p121.f90:86-104, p123.f90:120-151).
 *   r0     : starting r (neq), the assembled load vector (penalty rows
 *            already set to store*val for p123, p123.f90:126-132)
 *   diag   : assembled diagonal (neq), penalty already added for fixed rows
 *   nfix, no_f (1-based eq), store : p123 fixed-freedom override
 *   stop   : OR_STOP_CHECON (checon_par, tol on max|dx|/max|x|) or
 *            OR_STOP_RELRES (sqrt((r,r)/(r0,r0)) <= tol)
 *   dot_order : OR_DOT_SEQ (DOT_PRODUCT at one rank), OR_DOT_THREADS
 *            (per-thread chunks), OR_DOT_DEVICE (the device library's order)
 * Outputs x (neq) = xnew at exit, *iters, and returns 1 if converged.
 * hist (nullable, length limit) receives max|dx|/max|x| (checon) or
 * ||r||/||r0|| (relres) per iteration. */
int or_pcg(int nels, int ntot, int neq, const int *g_g, const double *storkm,
           const double *r0, const double *diag, int nfix, const int *no_f,
           const double *store, double tol, int limit, int stop, int nthreads,
           int dot_order, double *x_out, int *iters_out, double *hist)
{
    int nt = omp_threads(nthreads);
    size_t nslot = (size_t)nels * ntot;
    double *r = (double *)malloc(sizeof(double) * neq);
    double *p = (double *)malloc(sizeof(double) * neq);
    double *d = (double *)malloc(sizeof(double) * neq);
    double *u = (double *)malloc(sizeof(double) * neq);
    double *x = (double *)calloc(neq, sizeof(double));
    double *xnew = (double *)calloc(neq, sizeof(double));
    double *dinv = (double *)malloc(sizeof(double) * neq);
    double *pmul = (double *)malloc(sizeof(double) * nslot);
    double *utemp = (double *)malloc(sizeof(double) * nslot);
    csr_t csr = build_csr(nels, ntot, neq, g_g);
    for (long i = 0; i < neq; ++i) {
        dinv[i] = 1.0 / diag[i];
        r[i] = r0[i];
        d[i] = dinv[i] * r[i];
        p[i] = d[i];
    }
    double rr0 = dot_mode(neq, r, r, nt, dot_order);
    int iters = 0, converged = 0;
    for (;;) {
        iters++;
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long k = 0; k < (long)nslot; ++k) pmul[k] = g_g[k] > 0 ? p[g_g[k] - 1] : 0.0;
        or_elem_matvec(nels, ntot, storkm, pmul, utemp, nt);
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long q = 0; q < neq; ++q) {
            double s = 0.0;
            for (long k = csr.ptr[q]; k < csr.ptr[q + 1]; ++k) s = s + utemp[csr.slot[k]];
            u[q] = s;
        }
        for (int i = 0; i < nfix; ++i) u[no_f[i] - 1] = p[no_f[i] - 1] * store[i];
        double up = dot_mode(neq, r, d, nt, dot_order);
        double alpha = up / dot_mode(neq, p, u, nt, dot_order);
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long i = 0; i < neq; ++i) {
            xnew[i] = x[i] + p[i] * alpha;
            r[i] = r[i] - u[i] * alpha;
            d[i] = dinv[i] * r[i];
        }
        double beta = dot_mode(neq, r, d, nt, dot_order) / up;
#pragma omp parallel for schedule(static) num_threads(nt)
        for (long i = 0; i < neq; ++i) p[i] = d[i] + p[i] * beta;
        double crit;
        if (stop == OR_STOP_CHECON) {
            double ml = 0.0, md = 0.0;
            for (long i = 0; i < neq; ++i) {
                double a = fabs(xnew[i]), dd = fabs(xnew[i] - x[i]);
                if (a > ml) ml = a;
                if (dd > md) md = dd;
                x[i] = xnew[i];
            }
            crit = md / ml;
        } else {
            memcpy(x, xnew, sizeof(double) * neq);
            crit = sqrt(dot_mode(neq, r, r, nt, dot_order) / rr0);
        }
        if (hist) hist[iters - 1] = crit;
        converged = crit <= tol;
        if (converged || iters == limit) break;
    }
    memcpy(x_out, xnew, sizeof(double) * neq);
    *iters_out = iters;
    free(r); free(p); free(d); free(u); free(x); free(xnew); free(dinv);
    free(pmul); free(utemp); free(csr.ptr); free(csr.slot);
    return converged;
}

/* Assembled diagonal of the element matrices (p121.f90:65-69). */
void or_diag(int nels, int ntot, int neq, const int *g_g, const double *storkm, double *diag)
{
    memset(diag, 0, sizeof(double) * neq);
    for (size_t e = 0; e < (size_t)nels; ++e)
        for (int i = 0; i < ntot; ++i) {
            int g = g_g[e * ntot + i];
            if (g > 0) diag[g - 1] = diag[g - 1] + storkm[e * ntot * ntot + IDX2(i, i, ntot)];
        }
}

/* K.x for a full vector (neq) via gather, element products, scatter. */
void or_apply(int nels, int ntot, int neq, const int *g_g, const double *storkm,
              const double *x, double *y, int nthreads)
{
    size_t nslot = (size_t)nels * ntot;
    double *pmul = (double *)malloc(sizeof(double) * nslot);
    double *ut = (double *)malloc(sizeof(double) * nslot);
    or_gather(nels, ntot, g_g, x, pmul);
    or_elem_matvec(nels, ntot, storkm, pmul, ut, nthreads);
    memset(y, 0, sizeof(double) * neq);
    or_scatter(nels, ntot, g_g, ut, y);
    free(pmul);
    free(ut);
}
