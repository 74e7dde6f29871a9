// p122's viscoplastic stress update (src/programs/5th_ed/p122/p122.f90:
// 164-199) for every element of an 8-node brick mesh, one element per
// thread, its Gauss points in SAMPLE's order:
//   eps = MATMUL(bee, eld) - evpt;  sigma = MATMUL(dee, eps);
//   stress = sigma + tensor;  INVAR;  MOCOUF -> f
//   final:  devp = stress
//   f >= 0: MOCOUQ, FORMM, flow = f*(m1*dq1 + m2*dq2 + m3*dq3),
//           erate = MATMUL(flow, stress), evp = erate*dt, evpt += evp,
//           devp = MATMUL(dee, evp)
//   f >= 0: bload += MATMUL(TRANSPOSE(bee), devp)*det*weights(i)
//   final:  tensor = stress
// `final` is p122's (plastic_converged .OR. plasiters == plasits).  Element
// routines: INVAR (new_library.f90:1813-1900, nst = 6), MOCOUF (:2364-2382),
// MOCOUQ (:2423-2460), FORMM (:85-198), BEEMAT (:918-1000), DEEMAT
// (:1604-1691).  Products and sums follow the reference's order; the
// trigonometric functions are the device library's, so results agree with
// a host evaluation to rounding, not bit for bit.
#include "pf_internal.h"
#include "pf_element.h"

namespace {

struct Mc {   // Mohr-Coulomb and elastic constants of the run
    double dee[6][6];
    double phi, c, psi, dt;
};

__device__ void invar6(const double s[6], double &sigm, double &dsbar, double &theta)
{
    const double sq3 = sqrt(3.0);
    const double s1 = s[0], s2 = s[1], s3 = s[2], s4 = s[3], s5 = s[4], s6 = s[5];
    sigm = (s1 + s2 + s3) / 3.0;
    const double d2 = ((s1 - s2) * (s1 - s2) + (s2 - s3) * (s2 - s3) + (s3 - s1) * (s3 - s1)) / 6.0 + s4 * s4 +
                      s5 * s5 + s6 * s6;
    const double ds1 = s1 - sigm, ds2 = s2 - sigm, ds3 = s3 - sigm;
    const double d3 = ds1 * ds2 * ds3 - ds1 * s5 * s5 - ds2 * s6 * s6 - ds3 * s4 * s4 + 2.0 * s4 * s5 * s6;
    dsbar = sq3 * sqrt(d2);
    if (dsbar < 1.e-10) {
        theta = 0.0;
    } else {
        const double r = sqrt(d2);
        double sine = -3.0 * sq3 * d3 / (2.0 * (r * r * r));
        if (sine > 1.0) sine = 1.0;
        if (sine < -1.0) sine = -1.0;
        theta = asin(sine) / 3.0;
    }
}

__device__ double mocouf(double phi, double c, double sigm, double dsbar, double theta)
{
    const double phir = phi * 4.0 * atan(1.0) / 180.0;
    const double snph = sin(phir), csph = cos(phir), csth = cos(theta), snth = sin(theta);
    return snph * sigm + dsbar * (csth / sqrt(3.0) - snth * snph / 3.0) - c * csph;
}

__device__ void mocouq(double psi, double dsbar, double theta, double &dq1, double &dq2, double &dq3)
{
    const double psir = psi * 4.0 * atan(1.0) / 180.0;
    const double snth = sin(theta), snps = sin(psir), sq3 = sqrt(3.0);
    dq1 = snps;
    if (fabs(snth) > 0.49) {
        const double c1 = snth < 0.0 ? -1.0 : 1.0;
        dq2 = (sq3 * 0.5 - c1 * snps * 0.5 / sq3) * sq3 * 0.5 / dsbar;
        dq3 = 0.0;
    } else {
        const double csth = cos(theta), cs3th = cos(3.0 * theta), tn3th = tan(3.0 * theta);
        const double tnth = snth / csth;
        dq2 = sq3 * csth / dsbar * ((1.0 + tnth * tn3th) + snps * (tn3th - tnth) / sq3) * 0.5;
        dq3 = 0.5 * 3.0 * (sq3 * snth + snps * csth) / (cs3th * dsbar * dsbar);
    }
}

// FORMM for nst = 6: m1, m2, m3 (6x6).
__device__ void formm6(const double s[6], double m1[6][6], double m2[6][6], double m3[6][6])
{
    const double sx = s[0], sy = s[1], sz = s[2], txy = s[3], tyz = s[4], tzx = s[5];
    const double sigm = (sx + sy + sz) / 3.0;
    const double dx = sx - sigm, dy = sy - sigm, dz = sz - sigm;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) m1[i][j] = m2[i][j] = m3[i][j] = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m1[i][j] = 1.0 / (3.0 * sigm);
    for (int i = 0; i < 3; ++i) {
        m2[i][i] = 2.0;
        m2[i + 3][i + 3] = 6.0;
    }
    m2[0][1] = -1.0; m2[0][2] = -1.0; m2[1][2] = -1.0;
    m3[0][0] = dx;  m3[0][1] = dz;  m3[0][2] = dy;  m3[0][3] = txy; m3[0][4] = -2.0 * tyz; m3[0][5] = tzx;
    m3[1][1] = dy;  m3[1][2] = dx;  m3[1][3] = txy; m3[1][4] = tyz; m3[1][5] = -2.0 * tzx;
    m3[2][2] = dz;  m3[2][3] = -2.0 * txy; m3[2][4] = tyz; m3[2][5] = tzx;
    m3[3][3] = -3.0 * dz; m3[3][4] = 3.0 * tzx; m3[3][5] = 3.0 * tyz;
    m3[4][4] = -3.0 * dx; m3[4][5] = 3.0 * txy;
    m3[5][5] = -3.0 * dy;
    for (int i = 0; i < 6; ++i)
        for (int j = i + 1; j < 6; ++j) {
            m1[j][i] = m1[i][j];
            m2[j][i] = m2[i][j];
            m3[j][i] = m3[i][j];
        }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            m1[i][j] = m1[i][j] / 3.0;
            m2[i][j] = m2[i][j] / 3.0;
            m3[i][j] = m3[i][j] / 3.0;
        }
}

__global__ __launch_bounds__(64) void k_plastic(const double *__restrict__ gc, const int *__restrict__ gg,
                                               const double *__restrict__ x, double *__restrict__ evpt,
                                               double *__restrict__ tensor, double *__restrict__ bload_out, Mc mc,
                                               int nels, int final_pass)
{
    const long iel = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (iel >= nels) return;
    const double *coord = gc + iel * 24;
    double eld[24];
    for (int i = 0; i < 24; ++i) {
        const int g = gg[iel * 24 + i];
        eld[i] = g ? x[g - 1] : 0.0;
    }
    double bload[24];
    for (int i = 0; i < 24; ++i) bload[i] = 0.0;
    constexpr int gs[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
                              {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    const double r3 = 1.0 / sqrt(3.0);
    for (int ip = 0; ip < 8; ++ip) {
        double deriv[3][8];
        const double det = hex8_deriv(coord, gs[ip][0] > 0 ? r3 : -r3, gs[ip][1] > 0 ? r3 : -r3,
                                      gs[ip][2] > 0 ? r3 : -r3, deriv);
        // eps = MATMUL(bee, eld), BEEMAT's entries in column order
        double eps[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int m = 0; m < 8; ++m) {
            const double dx = deriv[0][m], dy = deriv[1][m], dz = deriv[2][m];
            const double ux = eld[3 * m], uy = eld[3 * m + 1], uz = eld[3 * m + 2];
            eps[0] = eps[0] + dx * ux;
            eps[1] = eps[1] + dy * uy;
            eps[2] = eps[2] + dz * uz;
            eps[3] = eps[3] + dy * ux;
            eps[3] = eps[3] + dx * uy;
            eps[4] = eps[4] + dz * uy;
            eps[4] = eps[4] + dy * uz;
            eps[5] = eps[5] + dz * ux;
            eps[5] = eps[5] + dx * uz;
        }
        double *ev = evpt + (iel * 8 + ip) * 6;
        double *te = tensor + (iel * 8 + ip) * 6;
        double stress[6];
        for (int r = 0; r < 6; ++r) eps[r] = eps[r] - ev[r];
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
            for (int t = 0; t < 6; ++t) s = s + mc.dee[r][t] * eps[t];
            stress[r] = s + te[r];
        }
        double sigm, dsbar, theta;
        invar6(stress, sigm, dsbar, theta);
        const double f = mocouf(mc.phi, mc.c, sigm, dsbar, theta);
        double devp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (final_pass) {
            for (int r = 0; r < 6; ++r) devp[r] = stress[r];
        } else if (f >= 0.0) {
            double dq1, dq2, dq3;
            mocouq(mc.psi, dsbar, theta, dq1, dq2, dq3);
            double m1[6][6], m2[6][6], m3[6][6];
            formm6(stress, m1, m2, m3);
            double evp[6];
            for (int a = 0; a < 6; ++a) {
                double s = 0.0;
                for (int b = 0; b < 6; ++b) {
                    const double flow = f * (m1[a][b] * dq1 + m2[a][b] * dq2 + m3[a][b] * dq3);
                    s = s + flow * stress[b];
                }
                evp[a] = s * mc.dt;
            }
            for (int a = 0; a < 6; ++a) ev[a] = ev[a] + evp[a];
            for (int a = 0; a < 6; ++a) {
                double s = 0.0;
                for (int t = 0; t < 6; ++t) s = s + mc.dee[a][t] * evp[t];
                devp[a] = s;
            }
        }
        if (f >= 0.0) {
            // eload = MATMUL(TRANSPOSE(bee), devp); bload += eload*det*weights(i)
            for (int m = 0; m < 8; ++m) {
                const double dx = deriv[0][m], dy = deriv[1][m], dz = deriv[2][m];
                const double ex = (dx * devp[0] + dy * devp[3]) + dz * devp[5];
                const double ey = (dy * devp[1] + dx * devp[3]) + dz * devp[4];
                const double ez = (dz * devp[2] + dy * devp[4]) + dx * devp[5];
                bload[3 * m] = bload[3 * m] + ex * det * 1.0;
                bload[3 * m + 1] = bload[3 * m + 1] + ey * det * 1.0;
                bload[3 * m + 2] = bload[3 * m + 2] + ez * det * 1.0;
            }
        }
        if (final_pass)
            for (int r = 0; r < 6; ++r) te[r] = stress[r];
    }
    for (int i = 0; i < 24; ++i) bload_out[iel * 24 + i] = bload[i];
}

}  // namespace

extern "C" int pf_plastic_update(pf_ctx *c, const double *x_dev, double *evpt_dev, double *tensor_dev,
                                 double *utemp_dev, double e, double v, double phi, double coh, double psi,
                                 double dt, int final_pass)
{
    if (!c || !x_dev || !evpt_dev || !tensor_dev || !utemp_dev) return PF_ERR_ARG;
    if (!c->g_coord || !c->g_g) return PF_ERR_STATE;
    if (c->nod != 8 || c->ntot != 24) return PF_ERR_ARG;
    if (!(v > -1.0 && v < 0.5)) return PF_ERR_ARG;
    PF_HIP(hipSetDevice(c->device));
    Mc mc{};
    // DEEMAT (nst = 6): dee0 * e/(2(1+v)vv)
    const double v2 = v / (1.0 - v), vv = (1.0 - 2.0 * v) / (1.0 - v) * 0.5;
    const double fe = e / (2.0 * (1.0 + v) * vv);
    for (int i = 0; i < 3; ++i) mc.dee[i][i] = 1.0 * fe;
    for (int i = 3; i < 6; ++i) mc.dee[i][i] = vv * fe;
    mc.dee[0][1] = mc.dee[1][0] = mc.dee[0][2] = mc.dee[2][0] = mc.dee[1][2] = mc.dee[2][1] = v2 * fe;
    mc.phi = phi;
    mc.c = coh;
    mc.psi = psi;
    mc.dt = dt;
    hipLaunchKernelGGL(k_plastic, dim3((c->nels + 63) / 64), dim3(64), 0, c->stream, c->g_coord, c->g_g, x_dev,
                       evpt_dev, tensor_dev, utemp_dev, mc, c->nels, final_pass);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

// u(j) = p(j)*store(j) (mode 1) or 0 (mode 0) on the rows of the fixed
// freedoms set by pf_fixed_set (p122.f90:141-146).
namespace {
__global__ __launch_bounds__(256) void k_fixed_rows(double *__restrict__ u, const double *__restrict__ p,
                                                   const double *__restrict__ fixstore, int neq, int mode)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < neq; q += gridDim.x * blockDim.x)
        if (fixstore[q + 1] != 0.0) u[q] = mode ? p[q] * fixstore[q + 1] : 0.0;
}
}  // namespace

extern "C" int pf_fixed_rows(pf_ctx *c, double *u_dev, const double *p_dev, int mode)
{
    if (!c || !u_dev || (mode == 1 && !p_dev) || (mode != 0 && mode != 1)) return PF_ERR_ARG;
    if (!c->nfix) return PF_OK;
    PF_HIP(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_fixed_rows, dim3(pf_grid_for(c->neq)), dim3(256), 0, c->stream, u_dev, p_dev, c->fixstore,
                       c->neq, mode);
    PF_HIP(hipGetLastError());
    return PF_OK;
}
