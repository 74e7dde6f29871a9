// Fused element assembly kernels (p121 loop A, p123 element loop).
//
// One 64-lane wavefront builds one element: SAMPLE's 2x2x2 Gauss points,
// SHAPE_DER at each point, the jacobian, its determinant and inverse, global
// derivatives, then the element matrix, written once to HBM: in the
// library's FP64 column-pair layout (pf_internal.h), or -- for FP32 storage
// -- straight into the packed symmetric hi/lo layout (hi = fl32(K), lo =
// fl32(K - hi) of the upper triangle) together with the element's FP64
// diagonal, so an FP32-stored problem never holds the FP64 matrices.
// Nothing but the element coordinates is read and nothing but the matrix is
// written.
//
// Reference: src/programs/5th_ed/p121/p121.f90:54-64 (elastic),
// src/programs/5th_ed/p123/p123.f90:70-84 (Laplace); element routines in
// src/modules/shared/new_library.f90 (SAMPLE :1406-1433, SHAPE_DER hex8
// :769-794, hex20 :865-896, BEEMAT :976-993, DEEMAT :1671-1686) and
// src/modules/mpi/maths.f90 (INVERT :525-547, DETERMINANT :615-619).
//
// Every quantity is formed with the reference's operation order (sums in
// ascending index order, each product and sum rounded, no fused
// multiply-add), skipping only products with the structural zeros of B and
// dee, which add exact zeros.  The element matrices therefore equal the ones
// a sequential evaluation of p121's MATMUL(MATMUL(TRANSPOSE(bee),dee),bee)
// *det*weights(i) gives, bit for bit.
#include "pf_internal.h"
#include "pf_element.h"

namespace {

constexpr int WAVES = 4;   // elements per block

// Per-wave LDS scratch for one element.
template <int NOD>
struct ElemLds {
    double coord[NOD * 3];      // (nod,3) column-major, as g_coord_pp(:,:,iel)
    double der[8][3][NOD];      // local derivatives per Gauss point
    double jac[8][9];           // jacobian (3x3 column-major), then its inverse
    double dw[8];               // det * weight
    double deriv[8][3][NOD];    // global derivatives
};

// Stages shared by both element types: Gauss points, local derivatives,
// jacobian, DETERMINANT/INVERT, global derivatives.  All four waves of the
// block call it (block-uniform barriers); `valid` masks elements past nels.
template <int NOD>
__device__ __forceinline__ void element_geometry(ElemLds<NOD> &L, const double *__restrict__ gc, int lane, bool valid)
{
    const double r3 = 1.0 / sqrt(3.0);
    if (valid)
        for (int k = lane; k < NOD * 3; k += 64) L.coord[k] = gc[k];
    // SAMPLE hexahedron nip=8 (new_library.f90:1406-1433): the signs of
    // (xi,eta,zeta) at point g, in the reference's order.
    constexpr int gs[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
                              {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    for (int k = lane; k < 8 * NOD; k += 64) {
        const int g = k / NOD, l = k % NOD;
        const double xi = gs[g][0] > 0 ? r3 : -r3;
        const double eta = gs[g][1] > 0 ? r3 : -r3;
        const double zeta = gs[g][2] > 0 ? r3 : -r3;
        double d3[3];
        shape_der3<NOD>(xi, eta, zeta, l, d3);
        L.der[g][0][l] = d3[0];
        L.der[g][1][l] = d3[1];
        L.der[g][2][l] = d3[2];
    }
    __syncthreads();
    // jac = MATMUL(der, coord): 72 entries (8 points x 3 x 3).
    for (int k = lane; k < 72; k += 64) {
        const int g = k / 9, ab = k % 9, a = ab % 3, b = ab / 3;
        double s = 0.0;
        for (int l = 0; l < NOD; ++l) s += L.der[g][a][l] * L.coord[l + NOD * b];
        L.jac[g][a + 3 * b] = s;
    }
    __syncthreads();
    if (lane < 8) {
        double *m = L.jac[lane];
#define M(a, b) m[((a)-1) + 3 * ((b)-1)]
        double det = M(1, 1) * (M(2, 2) * M(3, 3) - M(3, 2) * M(2, 3));
        det = det - M(1, 2) * (M(2, 1) * M(3, 3) - M(3, 1) * M(2, 3));
        det = det + M(1, 3) * (M(2, 1) * M(3, 2) - M(3, 1) * M(2, 2));
        const double j11 = M(2, 2) * M(3, 3) - M(3, 2) * M(2, 3);
        const double j21 = -M(2, 1) * M(3, 3) + M(3, 1) * M(2, 3);
        const double j31 = M(2, 1) * M(3, 2) - M(3, 1) * M(2, 2);
        const double j12 = -M(1, 2) * M(3, 3) + M(3, 2) * M(1, 3);
        const double j22 = M(1, 1) * M(3, 3) - M(3, 1) * M(1, 3);
        const double j32 = -M(1, 1) * M(3, 2) + M(3, 1) * M(1, 2);
        const double j13 = M(1, 2) * M(2, 3) - M(2, 2) * M(1, 3);
        const double j23 = -M(1, 1) * M(2, 3) + M(2, 1) * M(1, 3);
        const double j33 = M(1, 1) * M(2, 2) - M(2, 1) * M(1, 2);
        M(1, 1) = j11 / det; M(1, 2) = j12 / det; M(1, 3) = j13 / det;
        M(2, 1) = j21 / det; M(2, 2) = j22 / det; M(2, 3) = j23 / det;
        M(3, 1) = j31 / det; M(3, 2) = j32 / det; M(3, 3) = j33 / det;
#undef M
        L.dw[lane] = det;   // SAMPLE weights are 1 for nip = 8
    }
    __syncthreads();
    // deriv = MATMUL(jac^-1, der)
    for (int k = lane; k < 8 * NOD; k += 64) {
        const int g = k / NOD, l = k % NOD;
        const double *ji = L.jac[g];
        const double d0 = L.der[g][0][l], d1 = L.der[g][1][l], d2 = L.der[g][2][l];
        for (int a = 0; a < 3; ++a) L.deriv[g][a][l] = ji[a] * d0 + ji[a + 3] * d1 + ji[a + 6] * d2;
    }
    __syncthreads();
}

// Elastic element.  Lane i is row i of K_e (node a = i/3, component ci).
// The entry (i,j) at a Gauss point is sum_k btd(i,k) B(k,j) with
// btd(i,k) = sum_l B(l,i) dee(l,k), evaluated in the reference's order
// (l then k ascending, every product rounded, no fused multiply-add); the
// zero entries of B and dee contribute exact zeros, so only the non-zero
// products are formed:
//   x dof: B(1,i)=dx B(4,i)=dy B(6,i)=dz   y dof: B(2,i)=dy B(4,i)=dx B(5,i)=dz
//   z dof: B(3,i)=dz B(5,i)=dy B(6,i)=dx   (BEEMAT ih=6)
// and km(i,j) += (s*det)*w with SAMPLE's w = 1.
// FP32 storage: entry (i, j), j >= i, of element e (the upper triangle the
// packed layout keeps) as hi and lo; the diagonal also in FP64 into dg(i,e)
// for the assembled Jacobi diagonal.
template <int NT>
__device__ __forceinline__ void store_packed(long e, int i, int j, double k, float *__restrict__ hi,
                                             float *__restrict__ lo, double *__restrict__ dg)
{
    if (j < i) return;
    const long t = pf_idx_sym_ij(e, NT, i, j);
    const float h = (float)k;
    hi[t] = h;
    lo[t] = (float)(k - (double)h);
    if (j == i) dg[e * NT + i] = k;
}

// Zero padding at the end of element e's packed block.
template <int NT>
__device__ __forceinline__ void zero_padding(long e, int lane, float *__restrict__ hi, float *__restrict__ lo)
{
    constexpr int used = NT * (NT + 1) / 2, pad = (int)pf_sym_per(NT) - used;
    if (lane < pad) {
        hi[e * pf_sym_per(NT) + used + lane] = 0.0f;
        lo[e * pf_sym_per(NT) + used + lane] = 0.0f;
    }
}

template <int NOD, bool S32>
__global__ __launch_bounds__(64 * WAVES) void k_assemble_elastic(
    const double *__restrict__ g_coord, const double *__restrict__ e_elem, double v2, double vv,
    double den, double e0, double *__restrict__ K, float *__restrict__ hi, float *__restrict__ lo,
    double *__restrict__ dg, int nels)
{
    constexpr int NTOT = 3 * NOD;
    __shared__ ElemLds<NOD> lds[WAVES];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long e = (long)blockIdx.x * WAVES + wave;
    const bool valid = e < nels;
    ElemLds<NOD> &L = lds[wave];
    element_geometry<NOD>(L, g_coord + (valid ? e : 0) * (NOD * 3), lane, valid);
    if (!valid) return;
    if constexpr (S32) zero_padding<NTOT>(e, lane, hi, lo);
    if (lane >= NTOT) return;
    // DEEMAT: dee = dee0 * (e / (2(1+v)vv)) with dee0 = 1 on the normal
    // diagonal, v2 off it, vv on the shear diagonal.
    const double f = (e_elem ? e_elem[e] : e0) / den;
    const double D11 = 1.0 * f, D12 = v2 * f, D44 = vv * f;
    const int a = lane / 3, ci = lane - 3 * (lane / 3);
    double bt[8][6];
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const double dx = L.deriv[g][0][a], dy = L.deriv[g][1][a], dz = L.deriv[g][2][a];
        if (ci == 0) {
            bt[g][0] = dx * D11; bt[g][1] = dx * D12; bt[g][2] = dx * D12;
            bt[g][3] = dy * D44; bt[g][4] = 0.0;      bt[g][5] = dz * D44;
        } else if (ci == 1) {
            bt[g][0] = dy * D12; bt[g][1] = dy * D11; bt[g][2] = dy * D12;
            bt[g][3] = dx * D44; bt[g][4] = dz * D44; bt[g][5] = 0.0;
        } else {
            bt[g][0] = dz * D12; bt[g][1] = dz * D12; bt[g][2] = dz * D11;
            bt[g][3] = 0.0;      bt[g][4] = dy * D44; bt[g][5] = dx * D44;
        }
    }
    // Two nodes (6 columns = 3 column pairs) per step so that every store
    // is one 16-byte column pair of the FP64 pair layout.
    double2 *Kp = reinterpret_cast<double2 *>(K + e * (long)(NTOT * NTOT)) + lane;
    for (int bb = 0; bb < NOD / 2; ++bb) {
        double o[6];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int b = 2 * bb + h;
            double o0 = 0.0, o1 = 0.0, o2 = 0.0;
#pragma unroll
            for (int g = 0; g < 8; ++g) {
                const double dx = L.deriv[g][0][b], dy = L.deriv[g][1][b], dz = L.deriv[g][2][b];
                const double det = L.dw[g];
                const double *t = bt[g];
                const double sx = (t[0] * dx + t[3] * dy) + t[5] * dz;   // column x dof of node b
                const double sy = (t[1] * dy + t[3] * dx) + t[4] * dz;   // column y dof
                const double sz = (t[2] * dz + t[4] * dy) + t[5] * dx;   // column z dof
                o0 = o0 + sx * det;
                o1 = o1 + sy * det;
                o2 = o2 + sz * det;
            }
            o[3 * h + 0] = o0;
            o[3 * h + 1] = o1;
            o[3 * h + 2] = o2;
        }
        if constexpr (S32) {
#pragma unroll
            for (int q = 0; q < 6; ++q) store_packed<NTOT>(e, lane, 6 * bb + q, o[q], hi, lo, dg);
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) Kp[(3 * bb + q) * NTOT] = make_double2(o[2 * q], o[2 * q + 1]);
        }
    }
}

// p123: kc{x,y,z} = sum_g d{x,y,z}_a d{x,y,z}_b det w; storkc = kx*kcx + ky*kcy + kz*kcz.
template <bool S32>
__global__ __launch_bounds__(64 * WAVES) void k_assemble_laplace8(
    const double *__restrict__ g_coord, const double *__restrict__ k_elem, double kx, double ky,
    double kz, double *__restrict__ K, float *__restrict__ hi, float *__restrict__ lo, double *__restrict__ dg,
    int nels)
{
    __shared__ ElemLds<8> lds[WAVES];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long e = (long)blockIdx.x * WAVES + wave;
    const bool valid = e < nels;
    ElemLds<8> &L = lds[wave];
    element_geometry<8>(L, g_coord + (valid ? e : 0) * 24, lane, valid);
    if (!valid) return;
    const int a = lane & 7, b = lane >> 3;   // row a, column b
    double cx = 0.0, cy = 0.0, cz = 0.0;
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const double w = L.dw[g];
        cx += L.deriv[g][0][a] * L.deriv[g][0][b] * w;
        cy += L.deriv[g][1][a] * L.deriv[g][1][b] * w;
        cz += L.deriv[g][2][a] * L.deriv[g][2][b] * w;
    }
    const double s = k_elem ? k_elem[e] : 1.0;
    const double k = cx * (kx * s) + cy * (ky * s) + cz * (kz * s);
    if constexpr (S32) store_packed<8>(e, a, b, k, hi, lo, dg);
    else K[pf_idx_pair(e, 8, a, b)] = k;
}

// p124's element loop (src/programs/5th_ed/p124/p124.f90:73-86) on an
// 8-node brick: per Gauss point kc += MATMUL(MATMUL(TRANSPOSE(deriv),kay),
// deriv)*det*weights(i) with kay = diag(kx,ky,kz), pm += MATMUL(TRANSPOSE(
// funny),funny)*det*weights(i)*rho*cp with funny = SHAPE_FUN
// (new_library.f90:416-420); then storka = pm + kc*theta*dtim and storkb =
// pm - kc*(1-theta)*dtim.  prop(5,np_types) = kx, ky, kz, rho, cp per
// element type (READ_MATERIAL).  Every product and sum in the reference's
// order: MATMUL(TRANSPOSE(deriv),kay) has one nonzero term per entry (the
// others add exact zeros), the weights of SAMPLE are 1.
__global__ __launch_bounds__(64 * WAVES) void k_assemble_transient8(
    const double *__restrict__ g_coord, const double *__restrict__ prop, const int *__restrict__ etype,
    double theta, double dtim, double *__restrict__ Ka, double *__restrict__ Kb, int nels)
{
    __shared__ ElemLds<8> lds[WAVES];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long e = (long)blockIdx.x * WAVES + wave;
    const bool valid = e < nels;
    ElemLds<8> &L = lds[wave];
    element_geometry<8>(L, g_coord + (valid ? e : 0) * 24, lane, valid);
    if (!valid) return;
    const double *pr = prop + 5 * (long)(etype[e] - 1);
    const double kay[3] = {pr[0], pr[1], pr[2]}, rho = pr[3], cp = pr[4];
    const int a = lane & 7, b = lane >> 3;   // row a, column b
    // SHAPE_FUN, 8-node brick: node -> (xi, eta, zeta) factors, - = (1 - s),
    // + = (s + 1), each 0.125 * f_xi * f_eta * f_zeta
    constexpr int sx[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    constexpr int sy[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    constexpr int sz[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    constexpr int gs[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
                              {-1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}};
    const double r3 = 1.0 / sqrt(3.0);
    double kc = 0.0, pm = 0.0;
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const double xi = gs[g][0] > 0 ? r3 : -r3, eta = gs[g][1] > 0 ? r3 : -r3, zeta = gs[g][2] > 0 ? r3 : -r3;
        const double xim = 1.0 - xi, etam = 1.0 - eta, zetam = 1.0 - zeta;
        const double xip = xi + 1.0, etap = eta + 1.0, zetap = zeta + 1.0;
        const double fa = 0.125 * (sx[a] < 0 ? xim : xip) * (sy[a] < 0 ? etam : etap) * (sz[a] < 0 ? zetam : zetap);
        const double fb = 0.125 * (sx[b] < 0 ? xim : xip) * (sy[b] < 0 ? etam : etap) * (sz[b] < 0 ? zetam : zetap);
        const double det = L.dw[g];
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) t = t + (L.deriv[g][m][a] * kay[m]) * L.deriv[g][m][b];
        kc = kc + t * det * 1.0;
        pm = pm + fa * fb * det * 1.0 * rho * cp;
    }
    Ka[pf_idx_pair(e, 8, a, b)] = pm + kc * theta * dtim;
    Kb[pf_idx_pair(e, 8, a, b)] = pm - kc * (1.0 - theta) * dtim;
}

// hi = fl32(K), lo = fl32(K - hi) in the packed symmetric layout (upper
// triangle): hi + lo carries ~48 significant bits of K.
__global__ void k_split_fp32(const double *__restrict__ K, float *__restrict__ hi, float *__restrict__ lo,
                             int nt, long n)
{
    const long per = pf_sym_per(nt);
    const int Q4 = 4 * pf_sym_quads(nt), RF = pf_sym_rem(nt), H = nt / 2;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long e = q / per;
        const int w = (int)(q - e * per);
        int i, d;
        if (w < Q4 * nt) {
            const int dq = w / (4 * nt), rem = w - dq * 4 * nt;
            i = rem >> 2;
            d = 4 * dq + (rem & 3);
        } else if (w < H * nt) {
            const int w2 = w - Q4 * nt;
            i = w2 / RF;
            d = Q4 + (w2 - i * RF);
        } else if (w < H * nt + H) {
            i = w - H * nt;
            d = H;
        } else {                        // padding
            hi[q] = 0.0f;
            lo[q] = 0.0f;
            continue;
        }
        int j = i + d;
        if (j >= nt) j -= nt;
        const double k = K[pf_idx_pair(e, nt, i < j ? i : j, i < j ? j : i)];
        const float h = (float)k;
        hi[q] = h;
        lo[q] = (float)(k - (double)h);
    }
}

// storkm_pp chunk (column-major, elements e0..e0+nel-1) -> FP64 pair layout.
__global__ void k_from_colmajor(const double *__restrict__ src, double *__restrict__ K, int nt, long e0, long nel)
{
    const long nn = (long)nt * nt, n = nel * nn;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long el = q / nn, w = q - el * nn;
        const int j = (int)(w / nt), i = (int)(w - (long)j * nt);
        K[pf_idx_pair(e0 + el, nt, i, j)] = src[q];
    }
}

// internal layout -> storkm_pp chunk (FP64 pairs, or packed symmetric hi + lo).
__global__ void k_to_colmajor(const double *__restrict__ K, const float *__restrict__ hi,
                              const float *__restrict__ lo, double *__restrict__ dst, int nt, long e0, long nel,
                              long kstride)
{
    const long nn = (long)nt * nt, n = nel * nn;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long el = q / nn, w = q - el * nn;
        const int j = (int)(w / nt), i = (int)(w - (long)j * nt);
        if (K) {
            dst[q] = K[pf_idx_pair((e0 + el) * kstride, nt, i, j)];
        } else {
            const long t = pf_idx_sym_ij(e0 + el, nt, i, j);
            dst[q] = (double)hi[t] + (double)lo[t];
        }
    }
}

// Assembled diagonal: diag(q) = sum over the slots of q, ascending element,
// of storkm(i,i,e) (p121.f90:65-69 then SCATTER).  diag has a leading 0 slot.
__global__ void k_diag(const int *__restrict__ ptr, const int *__restrict__ slot, const double *__restrict__ K,
                       int ntot, long kstride, double *__restrict__ diag, int neq)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < neq; q += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = ptr[q]; k < ptr[q + 1]; ++k) {
            const long sl = slot[k];
            const long e = sl / ntot;
            const int i = (int)(sl - e * ntot);
            s = __dadd_rn(s, K[pf_idx_pair(e * kstride, ntot, i, i)]);
        }
        diag[q + 1] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) diag[0] = 0.0;
}

// The same diagonal from the FP64 element diagonals dg(ntot,nels) that the
// FP32-storage assembly writes.
__global__ void k_diag_elem(const int *__restrict__ ptr, const int *__restrict__ slot,
                            const double *__restrict__ dg, double *__restrict__ diag, int neq)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < neq; q += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = ptr[q]; k < ptr[q + 1]; ++k) s = __dadd_rn(s, dg[slot[k]]);
        diag[q + 1] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) diag[0] = 0.0;
}

}  // namespace

// FP64 storage writes K64; FP32 storage writes K32hi/K32lo and the element
// diagonals into dg (ntot*nels doubles).
int pf_launch_assemble_elastic(pf_ctx *c, double v2, double vv, double den, double e0, const double *e_elem,
                               double *dg)
{
    const int grid = (c->nels + WAVES - 1) / WAVES;
    const bool s32 = c->store == PF_STORE_FP32;
    if (s32 && (!c->K32hi || !c->K32lo || !dg)) return PF_ERR_STATE;
    if (c->nod == 20 && s32)
        hipLaunchKernelGGL((k_assemble_elastic<20, true>), dim3(grid), dim3(64 * WAVES), 0, c->stream, c->g_coord,
                           e_elem, v2, vv, den, e0, (double *)nullptr, c->K32hi, c->K32lo, dg, c->nels);
    else if (c->nod == 20)
        hipLaunchKernelGGL((k_assemble_elastic<20, false>), dim3(grid), dim3(64 * WAVES), 0, c->stream, c->g_coord,
                           e_elem, v2, vv, den, e0, c->K64, (float *)nullptr, (float *)nullptr, (double *)nullptr,
                           c->nels);
    else if (c->nod == 8 && s32)
        hipLaunchKernelGGL((k_assemble_elastic<8, true>), dim3(grid), dim3(64 * WAVES), 0, c->stream, c->g_coord,
                           e_elem, v2, vv, den, e0, (double *)nullptr, c->K32hi, c->K32lo, dg, c->nels);
    else if (c->nod == 8)
        hipLaunchKernelGGL((k_assemble_elastic<8, false>), dim3(grid), dim3(64 * WAVES), 0, c->stream, c->g_coord,
                           e_elem, v2, vv, den, e0, c->K64, (float *)nullptr, (float *)nullptr, (double *)nullptr,
                           c->nels);
    else
        return PF_ERR_ARG;
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_assemble_laplace(pf_ctx *c, double kx, double ky, double kz, const double *k_elem, double *dg)
{
    if (c->nod != 8) return PF_ERR_ARG;
    const int grid = (c->nels + WAVES - 1) / WAVES;
    if (c->store == PF_STORE_FP32) {
        if (!c->K32hi || !c->K32lo || !dg) return PF_ERR_STATE;
        hipLaunchKernelGGL(k_assemble_laplace8<true>, dim3(grid), dim3(64 * WAVES), 0, c->stream, c->g_coord,
                           k_elem, kx, ky, kz, (double *)nullptr, c->K32hi, c->K32lo, dg, c->nels);
    } else {
        hipLaunchKernelGGL(k_assemble_laplace8<false>, dim3(grid), dim3(64 * WAVES), 0, c->stream, c->g_coord,
                           k_elem, kx, ky, kz, c->K64, (float *)nullptr, (float *)nullptr, (double *)nullptr,
                           c->nels);
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_diag_elem(pf_ctx *c, const double *dg, double *diag)
{
    const int grid = (c->neq + 255) / 256 < 8192 ? (c->neq + 255) / 256 : 8192;
    hipLaunchKernelGGL(k_diag_elem, dim3(grid), dim3(256), 0, c->stream, c->csr_ptr, c->csr_slot, dg, diag, c->neq);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_assemble_transient(pf_ctx *c, double theta, double dtim, const double *prop, const int *etype)
{
    if (c->nod != 8) return PF_ERR_ARG;
    const int grid = (c->nels + WAVES - 1) / WAVES;
    hipLaunchKernelGGL(k_assemble_transient8, dim3(grid), dim3(64 * WAVES), 0, c->stream, c->g_coord, prop, etype,
                       theta, dtim, c->K64, c->K64b, c->nels);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_split_fp32(pf_ctx *c)
{
    const long n = pf_sym_per(c->ntot) * c->nels;
    hipLaunchKernelGGL(k_split_fp32, dim3(8192), dim3(256), 0, c->stream, c->K64, c->K32hi, c->K32lo, c->ntot, n);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_from_colmajor(pf_ctx *c, const double *src, long e0, long nel)
{
    hipLaunchKernelGGL(k_from_colmajor, dim3(8192), dim3(256), 0, c->stream, src, c->K64, c->ntot, e0, nel);
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_to_colmajor(pf_ctx *c, double *dst, long e0, long nel)
{
    hipLaunchKernelGGL(k_to_colmajor, dim3(8192), dim3(256), 0, c->stream, c->K64, c->K32hi, c->K32lo, dst,
                       c->ntot, e0, nel, pf_k64_estride(c));
    PF_HIP(hipGetLastError());
    return PF_OK;
}

int pf_launch_diag(pf_ctx *c, double *diag)
{
    const int grid = (c->neq + 255) / 256 < 8192 ? (c->neq + 255) / 256 : 8192;
    hipLaunchKernelGGL(k_diag, dim3(grid), dim3(256), 0, c->stream, c->csr_ptr, c->csr_slot, c->K64, c->ntot,
                       pf_k64_estride(c), diag, c->neq);
    PF_HIP(hipGetLastError());
    return PF_OK;
}
