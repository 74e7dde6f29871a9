// Element routines shared by the device kernels (one element or one Gauss
// point per thread or per wavefront).
#pragma once

#include <hip/hip_runtime.h>

// SHAPE_DER for hex8 (nod = 8) and hex20 (nod = 20) at (xi, eta, zeta):
// derivative k (0..2) of shape function l (0-based).
template <int NOD>
static __device__ __forceinline__ void shape_der3(double xi, double eta, double zeta, int l, double out[3])
{
    if constexpr (NOD == 8) {
        const double etam = 1.0 - eta, xim = 1.0 - xi, zetam = 1.0 - zeta;
        const double etap = eta + 1.0, xip = xi + 1.0, zetap = zeta + 1.0;
        // Table of SHAPE_DER case(8): der(k,l) = s * f1 * f2.
        switch (l) {
        case 0: out[0] = -0.125 * etam * zetam; out[1] = -0.125 * xim * zetam; out[2] = -0.125 * xim * etam; break;
        case 1: out[0] = -0.125 * etam * zetap; out[1] = -0.125 * xim * zetap; out[2] = 0.125 * xim * etam; break;
        case 2: out[0] = 0.125 * etam * zetap;  out[1] = -0.125 * xip * zetap; out[2] = 0.125 * xip * etam; break;
        case 3: out[0] = 0.125 * etam * zetam;  out[1] = -0.125 * xip * zetam; out[2] = -0.125 * xip * etam; break;
        case 4: out[0] = -0.125 * etap * zetam; out[1] = 0.125 * xim * zetam;  out[2] = -0.125 * xim * etap; break;
        case 5: out[0] = -0.125 * etap * zetap; out[1] = 0.125 * xim * zetap;  out[2] = 0.125 * xim * etap; break;
        case 6: out[0] = 0.125 * etap * zetap;  out[1] = 0.125 * xip * zetap;  out[2] = 0.125 * xip * etap; break;
        default: out[0] = 0.125 * etap * zetam; out[1] = 0.125 * xip * zetam; out[2] = -0.125 * xip * etap; break;
        }
    } else {
        // Node sign tables of SHAPE_DER case(20).
        constexpr int xii[20] = {-1, -1, -1, 0, 1, 1, 1, 0, -1, -1, 1, 1, -1, -1, -1, 0, 1, 1, 1, 0};
        constexpr int etai[20] = {-1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1};
        constexpr int zetai[20] = {-1, 0, 1, 1, 1, 0, -1, -1, -1, 1, 1, -1, -1, 0, 1, 1, 1, 0, -1, -1};
        const double sx = xii[l], sy = etai[l], sz = zetai[l];
        const double xi0 = xi * sx, eta0 = eta * sy, zeta0 = zeta * sz;
        const int n = l + 1;
        if (n == 4 || n == 8 || n == 16 || n == 20) {
            out[0] = -.5 * xi * (1. + eta0) * (1. + zeta0);
            out[1] = .25 * sy * (1. - xi * xi) * (1. + zeta0);
            out[2] = .25 * sz * (1. - xi * xi) * (1. + eta0);
        } else if (n >= 9 && n <= 12) {
            out[0] = .25 * sx * (1. - eta * eta) * (1. + zeta0);
            out[1] = -.5 * eta * (1. + xi0) * (1. + zeta0);
            out[2] = .25 * sz * (1. + xi0) * (1. - eta * eta);
        } else if (n == 2 || n == 6 || n == 14 || n == 18) {
            out[0] = .25 * sx * (1. + eta0) * (1. - zeta * zeta);
            out[1] = .25 * sy * (1. + xi0) * (1. - zeta * zeta);
            out[2] = -.5 * zeta * (1. + xi0) * (1. + eta0);
        } else {
            out[0] = .125 * sx * (1. + eta0) * (1. + zeta0) * (2. * xi0 + eta0 + zeta0 - 1.);
            out[1] = .125 * sy * (1. + xi0) * (1. + zeta0) * (xi0 + 2. * eta0 + zeta0 - 1.);
            out[2] = .125 * sz * (1. + xi0) * (1. + eta0) * (xi0 + eta0 + 2. * zeta0 - 1.);
        }
    }
}

// One Gauss point of an 8-node brick, as p121/p122 evaluate it
// (p122.f90:170-173): der = SHAPE_DER at (xi, eta, zeta), jac =
// MATMUL(der, coord), det = DETERMINANT(jac) (maths.f90:615-619), INVERT
// (maths.f90:525-547), deriv = MATMUL(jac, der).  coord: (8,3) column-major.
static __device__ __forceinline__ double hex8_deriv(const double *coord, double xi, double eta, double zeta,
                                                   double deriv[3][8])
{
    double der[3][8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        double d3[3];
        shape_der3<8>(xi, eta, zeta, l, d3);
        der[0][l] = d3[0];
        der[1][l] = d3[1];
        der[2][l] = d3[2];
    }
    double m[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int l = 0; l < 8; ++l) s = s + der[a][l] * coord[l + 8 * b];
            m[a][b] = s;
        }
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
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) ji[a][b] = ji[a][b] / det;
#pragma unroll
    for (int l = 0; l < 8; ++l)
#pragma unroll
        for (int a = 0; a < 3; ++a) deriv[a][l] = ji[a][0] * der[0][l] + ji[a][1] * der[1][l] + ji[a][2] * der[2][l];
    return det;
}
