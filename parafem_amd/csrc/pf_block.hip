// Multi-RHS block EBE-PCG: up to 64 load cases solved together.
//
// The element product for all right-hand sides, U_e = K_e P_e with P_e the
// (ntot x 64) gathered block, runs on the matrix cores: one wavefront per
// element accumulates 16x16 output tiles with v_mfma_f32_16x16x4f32 (inner
// solves) or v_mfma_f64_16x16x4f64 (true residuals).  Tile operand layout of
// these instructions (lane l of 64): A(16x4) element [l%16][l/16], B(4x16)
// element [l/16][l%16]; D(16x16) register v holds [4(l/16)+v][l%16] for
// the FP32 instruction and [4v+l/16][l%16] for the FP64 one.
// Element matrices are stored pre-arranged as A operands: for tile row mt
// and k-step quad ktq, lane l reads four consecutive values (k-steps
// 4ktq..4ktq+3) -- one 16-byte (FP32) or 32-byte (FP64) load per lane,
// 1 KB / 2 KB per wavefront, coalesced.  The FP32 matrices keep only the
// 16x16 tiles on and above the diagonal (upper triangle, mirrored inside
// the diagonal tiles): the product uses each off-diagonal tile as stored for
// its own rows and, transposed through LDS, for the mirrored rows.  Block vectors keep column c at
// position pf_blk_pos(c), so the B operand of lane l for all four column
// tiles is one vector load from the gathered row.
//
// Solver: one FP32 block PCG (vectors FP32, dots accumulated in FP64, every
// column with its own alpha and beta) with residual replacement in FP64:
// when a column's recurrence residual has fallen by delta since its last
// replacement (or below tol), the FP32 solution is added into the FP64 one,
// r = f - K x is recomputed with the FP64 element matrices, and the FP32
// recurrence continues from fl32(r) with the same search directions.  A
// column stops once its true relative residual is below tol.
//
// The reference has no multi-RHS solve; the single-RHS iteration it batches
// is p121.f90:86-104 and its identical-K_e variant PCG_KM
// (src/modules/mpi/maths.f90:1152-1323, utemp_pp = MATMUL(km, pmul_pp), the
// same matrix-times-many-vectors shape).  The Jacobi preconditioner is
// p121.f90:65-69,86.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pf_internal.h"

namespace {

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d4_t __attribute__((ext_vector_type(4)));

template <typename T> struct V4;
template <> struct V4<float> { typedef f4_t type; };
template <> struct V4<double> { typedef d4_t type; };

__device__ __forceinline__ f4_t mfma(float a, float b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ d4_t mfma(double a, double b, d4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Tile geometry of an ntot x ntot element matrix.
template <int NT> struct Tiles {
    static constexpr int MT = (NT + 15) / 16;      // 16-row output tiles
    static constexpr int KT = (NT + 3) / 4;        // 4-wide k-steps
    static constexpr int KQ = (KT + 3) / 4;        // k-step quads (one vector load)
    static constexpr int PER = MT * KQ * 256;      // stored values per element
};

// K64 (column pairs) -> A-operand tiles [e][mt][kq][lane][j], zero padded.
template <int NT, typename T>
__global__ void k_blk_pack(const double *__restrict__ K, long kstride, T *__restrict__ Kt, long nels)
{
    using TL = Tiles<NT>;
    const long n = nels * TL::PER;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long e = q / TL::PER;
        const int w = (int)(q - e * TL::PER);
        const int j = w & 3, l = (w >> 2) & 63, kq = (w >> 8) % TL::KQ, mt = (w >> 8) / TL::KQ;
        const int row = 16 * mt + (l & 15), col = 4 * (4 * kq + j) + (l >> 4);
        Kt[q] = (row < NT && col < NT) ? (T)K[pf_idx_pair(e * kstride, NT, row, col)] : (T)0;
    }
}

// U_e = K_e P_e for every element, all PF_BLK columns, from full tiles (the
// FP64 product of the true residuals); MP output tile rows per pass.
template <int NT, typename T, int MP>
__global__ __launch_bounds__(256) void k_blk_ebe(const T *__restrict__ Kt, long kstride, const int *__restrict__ gg,
                                                const T *__restrict__ P, T *__restrict__ Ut, int nels,
                                                const pf_blk_state *__restrict__ st)
{
    using TL = Tiles<NT>;
    using T4 = typename V4<T>::type;
    if (st && st->done) return;
    const int l = threadIdx.x & 63;
    const long e = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (e >= nels) return;                       // whole wavefronts only; no block barriers below
    const int cl = l & 15, rq = l >> 4;
    const int gl = l < NT ? gg[e * NT + l] : 0;
    const T *Ke = Kt + e * kstride * TL::PER;
#pragma unroll
    for (int m0 = 0; m0 < TL::MT; m0 += MP) {
        T4 acc[MP][4];
#pragma unroll
        for (int m = 0; m < MP; ++m)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[m][nt] = T4{0, 0, 0, 0};
#pragma unroll
        for (int kq = 0; kq < TL::KQ; ++kq) {
            T4 a[MP];
#pragma unroll
            for (int m = 0; m < MP; ++m)
                a[m] = (m0 + m < TL::MT)
                           ? __builtin_nontemporal_load(reinterpret_cast<const T4 *>(Ke) + ((m0 + m) * TL::KQ + kq) * 64 + l)
                           : T4{0, 0, 0, 0};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int kt = 4 * kq + j;
                if (kt >= TL::KT) break;
                const int row = 4 * kt + rq;
                const int g = __shfl(gl, row < NT ? row : 0, 64);
                T4 b = *reinterpret_cast<const T4 *>(P + (long)g * PF_BLK + 4 * cl);
                if (row >= NT) b = T4{0, 0, 0, 0};
#pragma unroll
                for (int m = 0; m < MP; ++m)
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt) acc[m][nt] = mfma(a[m][j], b[nt], acc[m][nt]);
            }
        }
#pragma unroll
        for (int m = 0; m < MP; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * (m0 + m) + (sizeof(T) == 4 ? 4 * rq + v : 4 * v + rq);
                if (m0 + m < TL::MT && row < NT)   // read back by the scatter only: non-temporal
                    __builtin_nontemporal_store(T4{acc[m][0][v], acc[m][1][v], acc[m][2][v], acc[m][3][v]},
                                                reinterpret_cast<T4 *>(Ut + (e * NT + row) * PF_BLK + 4 * cl));
            }
    }
}

// Symmetric FP32 tiles: the MT(MT+1)/2 tiles (a, b), a <= b, row-major over
// the upper triangle of tiles; tile (a, b) lane l holds S[l%16][4j + l/16],
// j = 0..3, with S = K[16a.., 16b..] and K(i, j) read from the upper
// triangle (K(min, max)), so the stored operator is exactly symmetric.
template <int NT> struct SymTiles {
    static constexpr int MT = Tiles<NT>::MT;
    static constexpr int NS = MT * (MT + 1) / 2;   // stored tiles
    static constexpr int PER = NS * 256;           // floats per element
};

template <int NT>
__global__ void k_blk_pack_sym(const double *__restrict__ K, long kstride, float *__restrict__ Ks, long nels)
{
    using ST = SymTiles<NT>;
    const long n = nels * ST::PER;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long e = q / ST::PER;
        const int w = (int)(q - e * ST::PER);
        const int j = w & 3, l = (w >> 2) & 63;
        int t = w >> 8, a = 0;
        while (t >= ST::MT - a) {                  // row a of the tile triangle holds MT - a tiles
            t -= ST::MT - a;
            ++a;
        }
        const int b = a + t;
        const int row = 16 * a + (l & 15), col = 16 * b + 4 * j + (l >> 4);
        const int lo = row < col ? row : col, hi = row < col ? col : row;
        Ks[q] = (hi < NT) ? (float)K[pf_idx_pair(e * kstride, NT, lo, hi)] : 0.0f;
    }
}

// U_e = K_e P_e in FP32 for every element, all PF_BLK columns, from the
// symmetric tiles.  One wavefront per element loads its tiles, writes the
// off-diagonal ones transposed into LDS, then walks the k-blocks in
// ascending order: for k-block kb it gathers the B operands of those four
// k-steps and adds into row tile r <= kb the stored tile (r, kb) and into
// row tile r > kb the transposed tile (kb, r), so every output tile sums its
// k-steps in ascending order.
template <int NT>
__global__ __launch_bounds__(256) void k_blk_ebe_sym(const float *__restrict__ Ks, long kstride,
                                                    const int *__restrict__ gg, const float *__restrict__ P,
                                                    float *__restrict__ Ut, int nels,
                                                    const pf_blk_state *__restrict__ st)
{
    using TL = Tiles<NT>;
    using ST = SymTiles<NT>;
    constexpr int MT = TL::MT, KT = TL::KT, NS = ST::NS, NO = NS - MT;
    __shared__ float tr[4][NO > 0 ? NO : 1][16][17];   // per wave, padded rows
    if (st && st->done) return;
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long e = (long)blockIdx.x * (blockDim.x >> 6) + w;
    if (e >= nels) return;                       // whole wavefronts only; no block barriers below
    const int cl = l & 15, rq = l >> 4;
    const int gl = l < NT ? gg[e * NT + l] : 0;
    const f4_t *Ke = reinterpret_cast<const f4_t *>(Ks) + e * kstride * (NS * 64);
    // tile (a, b), a <= b, sits at index a*MT - a(a-1)/2 + (b - a); off-diagonal
    // tile (a, b) at LDS slot a*(MT-1) - a(a-1)/2 + (b - a - 1)
    auto tix = [](int a, int b) { return a * MT - a * (a - 1) / 2 + (b - a); };
    auto oix = [](int a, int b) { return a * (MT - 1) - a * (a - 1) / 2 + (b - a - 1); };
    f4_t s[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) s[t] = __builtin_nontemporal_load(Ke + t * 64 + l);
    if constexpr (NO > 0) {
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int bb = a + 1; bb < MT; ++bb)
#pragma unroll
                for (int j = 0; j < 4; ++j) tr[w][oix(a, bb)][l & 15][4 * j + (l >> 4)] = s[tix(a, bb)][j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    f4_t acc[MT][4];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[m][nt] = f4_t{0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < MT; ++kb) {
        f4_t b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = 16 * kb + 4 * j + rq;
            const int g = __shfl(gl, row < NT ? row : 0, 64);
            b[j] = row < NT ? *reinterpret_cast<const f4_t *>(P + (long)g * PF_BLK + 4 * cl) : f4_t{0, 0, 0, 0};
        }
#pragma unroll
        for (int r = 0; r < MT; ++r)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (4 * kb + j >= KT) break;
                const float av = r <= kb ? s[tix(r, kb)][j] : tr[w][oix(kb, r)][4 * j + (l >> 4)][l & 15];
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) acc[r][nt] = mfma(av, b[j][nt], acc[r][nt]);
            }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 16 * m + 4 * rq + v;
            if (row < NT)   // read back by the scatter only: non-temporal
                __builtin_nontemporal_store(f4_t{acc[m][0][v], acc[m][1][v], acc[m][2][v], acc[m][3][v]},
                                            reinterpret_cast<f4_t *>(Ut + (e * NT + row) * PF_BLK + 4 * cl));
        }
}

// Per-column reduction of 16 rows x 4 waves of a block into partials[q][blk][pos].
__device__ void blk_block_partials(double (&v)[4], double *partials)
{
    __shared__ double ws[4][PF_BLK];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double s = v[k];
        s = __dadd_rn(s, __shfl_down(s, 32, 64));
        s = __dadd_rn(s, __shfl_down(s, 16, 64));
        v[k] = s;
    }
    if (lane < 16)
#pragma unroll
        for (int k = 0; k < 4; ++k) ws[wave][4 * lane + k] = v[k];
    __syncthreads();
    if (threadIdx.x < PF_BLK) {
        double s = ws[0][threadIdx.x];
        for (int w = 1; w < 4; ++w) s = __dadd_rn(s, ws[w][threadIdx.x]);
        partials[(long)blockIdx.x * PF_BLK + threadIdx.x] = s;
    }
    __syncthreads();   // ws is reused by the next call in the same kernel
}

// U(q) = sum over q's element slots, ascending element, of Ut(slot); 16
// threads per equation row (four columns each).  DOT: partials of P.U.
template <typename T, bool DOT>
__global__ __launch_bounds__(256) void k_blk_scatter(const int *__restrict__ ptr, const int *__restrict__ slot,
                                                    const T *__restrict__ Ut, const T *__restrict__ P,
                                                    T *__restrict__ U, int neq, double *partials,
                                                    const pf_blk_state *st)
{
    using T4 = typename V4<T>::type;
    if (st && st->done) return;
    const int cl = threadIdx.x & 15, rl = threadIdx.x >> 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = blockIdx.x * 16 + rl; q < neq; q += gridDim.x * 16) {
        // the slot indices of up to 8 slots, then their rows, are loaded
        // before the sums (in slot order); longer lists continue in a loop
        const int a = ptr[q], b = ptr[q + 1];
        int sl[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) sl[j] = a + j < b ? slot[a + j] : -1;
        T4 t[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            t[j] = sl[j] >= 0 ? *reinterpret_cast<const T4 *>(Ut + (long)sl[j] * PF_BLK + 4 * cl) : T4{0, 0, 0, 0};
        T4 s = T4{0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (sl[j] >= 0) s = s + t[j];
        for (int k = a + 8; k < b; ++k) s = s + *reinterpret_cast<const T4 *>(Ut + (long)slot[k] * PF_BLK + 4 * cl);
        *reinterpret_cast<T4 *>(U + (long)(q + 1) * PF_BLK + 4 * cl) = s;
        if (DOT) {
            const T4 p = *reinterpret_cast<const T4 *>(P + (long)(q + 1) * PF_BLK + 4 * cl);
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = __dadd_rn(acc[k], __dmul_rn((double)p[k], (double)s[k]));
        }
    }
    if (DOT) blk_block_partials(acc, partials);
}

// Folds G block partials of quantity q per column (fixed order) -> out[pos].
__device__ void blk_fold(const double *partials, int G, double *out)
{
    __shared__ double ws[4][PF_BLK];
    const int col = threadIdx.x & 63, part = threadIdx.x >> 6;
    // thread (col, part) adds blocks part, part+4, ... in that order, the
    // loads of a batch of 16 issued before the dependent additions
    double s = 0.0;
    for (int b0 = part; b0 < G; b0 += 64) {
        double t[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int b = b0 + 4 * u;
            t[u] = b < G ? partials[(long)b * PF_BLK + col] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (b0 + 4 * u < G) s = __dadd_rn(s, t[u]);
    }
    ws[part][col] = s;
    __syncthreads();
    if (threadIdx.x < PF_BLK) {
        double t = ws[0][col];
        for (int p = 1; p < 4; ++p) t = __dadd_rn(t, ws[p][col]);
        out[col] = t;
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_blk_fold_alpha(const double *partials, int G, pf_blk_state *st)
{
    if (st->done) {
        if (threadIdx.x == 0) st->xlast = 0;   // the last p update has run
        return;
    }
    __shared__ double pu[PF_BLK];
    blk_fold(partials, G, pu);
    if (threadIdx.x < PF_BLK) {
        const int c = threadIdx.x;
        const double a = st->frozen[c] || pu[c] == 0.0 ? 0.0 : st->up[c] / pu[c];
        st->alpha[c] = isfinite(a) ? a : 0.0;
    }
}

// r -= alpha u; partials of (r, dinv r) and (r,r).  The preconditioned
// residual d = dinv r is not stored, and x += alpha p waits for
// k_blk_pupdate, which reads p anyway: one vector pass fewer.
__global__ __launch_bounds__(256) void k_blk_update(float *__restrict__ R, const float *__restrict__ U,
                                                   const float *__restrict__ dinv, int neq, double *partials,
                                                   const pf_blk_state *st)
{
    if (st->done) return;
    const int cl = threadIdx.x & 15, rl = threadIdx.x >> 4;
    double rd[4] = {0, 0, 0, 0}, rr[4] = {0, 0, 0, 0};
    float a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = (float)st->alpha[4 * cl + k];
    const int qc = (neq + (int)gridDim.x - 1) / (int)gridDim.x;   // rows per block, one contiguous range
    const int qe = ((int)blockIdx.x + 1) * qc < neq ? ((int)blockIdx.x + 1) * qc : neq;
    for (int q = (int)blockIdx.x * qc + rl; q < qe; q += 16) {
        const long o = (long)(q + 1) * PF_BLK + 4 * cl;
        f4_t r = *reinterpret_cast<const f4_t *>(R + o);
        const f4_t u = *reinterpret_cast<const f4_t *>(U + o);
        const float di = dinv[q + 1];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            r[k] = __fmaf_rn(-a[k], u[k], r[k]);
            const float d = di * r[k];
            rd[k] = __dadd_rn(rd[k], __dmul_rn((double)r[k], (double)d));
            rr[k] = __dadd_rn(rr[k], __dmul_rn((double)r[k], (double)r[k]));
        }
        *reinterpret_cast<f4_t *>(R + o) = r;
    }
    blk_block_partials(rd, partials);
    blk_block_partials(rr, partials + (long)PF_BLK_MAXG * PF_BLK);
}

// beta = (r,d)/up per active column; a column whose recurrence residual fell
// by delta since its last replacement, or below tol ||f||, asks for a
// residual replacement (trig); done at the iteration limit.
__global__ __launch_bounds__(256) void k_blk_fold_update(const double *partials, int G, pf_blk_state *st)
{
    if (st->done) return;
    __shared__ double rd[PF_BLK], rr[PF_BLK];
    __shared__ int trig;
    blk_fold(partials, G, rd);
    blk_fold(partials + (long)PF_BLK_MAXG * PF_BLK, G, rr);
    if (threadIdx.x == 0) trig = 0;
    __syncthreads();
    if (threadIdx.x < PF_BLK) {
        const int c = threadIdx.x;
        if (!st->frozen[c]) {
            const double b = st->up[c] == 0.0 ? 0.0 : rd[c] / st->up[c];
            st->beta[c] = isfinite(b) ? b : 0.0;
            st->up[c] = rd[c];
            st->rr[c] = rr[c];
            if (!(rr[c] > st->delta2 * st->rr_ref[c]) || !(rr[c] > st->tol2 * st->ff[c])) atomicAdd(&trig, 1);
        } else {
            st->beta[c] = 0.0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        st->iters += 1;
        if (trig) st->trig = 1;
        if (st->iters >= st->limit) {
            st->done = 1;
            st->xlast = 1;
        }
    }
}

// x += alpha p, then p = dinv r + beta p (only the x update on the
// iteration that set done: the solve ends there)
__global__ __launch_bounds__(256) void k_blk_pupdate(float *__restrict__ X, float *__restrict__ P,
                                                    const float *__restrict__ R, const float *__restrict__ dinv,
                                                    int neq, const pf_blk_state *st)
{
    if (st->done && !st->xlast) return;
    const bool last = st->done;
    const int cl = threadIdx.x & 15, rl = threadIdx.x >> 4;
    float a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = (float)st->alpha[4 * cl + k];
        b[k] = (float)st->beta[4 * cl + k];
    }
    const int qc = (neq + (int)gridDim.x - 1) / (int)gridDim.x;   // rows per block, one contiguous range
    const int qe = ((int)blockIdx.x + 1) * qc < neq ? ((int)blockIdx.x + 1) * qc : neq;
    for (int q = (int)blockIdx.x * qc + rl; q < qe; q += 16) {
        const long o = (long)(q + 1) * PF_BLK + 4 * cl;
        f4_t p = *reinterpret_cast<const f4_t *>(P + o), x = *reinterpret_cast<const f4_t *>(X + o);
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = __fmaf_rn(a[k], p[k], x[k]);
        *reinterpret_cast<f4_t *>(X + o) = x;
        if (last) continue;
        const f4_t r = *reinterpret_cast<const f4_t *>(R + o);
        const float di = dinv[q + 1];
#pragma unroll
        for (int k = 0; k < 4; ++k) p[k] = __fmaf_rn(b[k], p[k], di * r[k]);
        __builtin_nontemporal_store(p, reinterpret_cast<f4_t *>(P + o));
    }
}

// R = fl32(R64) and partials (r, dinv r), (r,r); START also X = 0 and
// P = dinv R (the first search directions).  Without START this is the
// residual replacement: x and p are kept.
template <bool START>
__global__ __launch_bounds__(256) void k_blk_inner_init(float *__restrict__ X, float *__restrict__ R,
                                                       float *__restrict__ P, const double *__restrict__ R64,
                                                       const float *__restrict__ dinv, int neq, double *partials)
{
    const int cl = threadIdx.x & 15, rl = threadIdx.x >> 4;
    double rd[4] = {0, 0, 0, 0}, rr[4] = {0, 0, 0, 0};
    const int qc = (neq + (int)gridDim.x - 1) / (int)gridDim.x;   // rows per block, one contiguous range
    const int qe = ((int)blockIdx.x + 1) * qc < neq ? ((int)blockIdx.x + 1) * qc : neq;
    for (int q = (int)blockIdx.x * qc + rl; q < qe; q += 16) {
        const long o = (long)(q + 1) * PF_BLK + 4 * cl;
        const float di = dinv[q + 1];
        f4_t r, d;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            r[k] = (float)R64[o + k];
            d[k] = di * r[k];
            rd[k] = __dadd_rn(rd[k], __dmul_rn((double)r[k], (double)d[k]));
            rr[k] = __dadd_rn(rr[k], __dmul_rn((double)r[k], (double)r[k]));
        }
        *reinterpret_cast<f4_t *>(R + o) = r;
        if (START) {
            *reinterpret_cast<f4_t *>(X + o) = f4_t{0.f, 0.f, 0.f, 0.f};
            *reinterpret_cast<f4_t *>(P + o) = d;
        }
    }
    blk_block_partials(rd, partials);
    blk_block_partials(rr, partials + (long)PF_BLK_MAXG * PF_BLK);
}

// After a start or a replacement: up = (r,d), rr = rr_ref = (r,r), no
// pending request.  Frozen flags and the tolerances are set by the host.
__global__ __launch_bounds__(256) void k_blk_fold_init(const double *partials, int G, pf_blk_state *st)
{
    __shared__ double rd[PF_BLK], rr[PF_BLK];
    blk_fold(partials, G, rd);
    blk_fold(partials + (long)PF_BLK_MAXG * PF_BLK, G, rr);
    if (threadIdx.x < PF_BLK) {
        const int c = threadIdx.x;
        st->up[c] = rd[c];
        st->rr[c] = rr[c];
        st->rr_ref[c] = rr[c];
        st->alpha[c] = st->beta[c] = 0.0;
    }
    if (threadIdx.x == 0) st->trig = 0;
}

// X64 += X (FP32 correction), then X = 0 when CLEAR.
template <bool CLEAR>
__global__ __launch_bounds__(256) void k_blk_correct(double *__restrict__ X64, float *__restrict__ X, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        X64[i + PF_BLK] = __dadd_rn(X64[i + PF_BLK], (double)X[i + PF_BLK]);
        if (CLEAR) X[i + PF_BLK] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_blk_residual(double *__restrict__ R64, const double *__restrict__ F64,
                                                     const double *__restrict__ U64, int neq, double *partials)
{
    const int cl = threadIdx.x & 15, rl = threadIdx.x >> 4;
    double rr[4] = {0, 0, 0, 0};
    const int qc = (neq + (int)gridDim.x - 1) / (int)gridDim.x;   // rows per block, one contiguous range
    const int qe = ((int)blockIdx.x + 1) * qc < neq ? ((int)blockIdx.x + 1) * qc : neq;
    for (int q = (int)blockIdx.x * qc + rl; q < qe; q += 16) {
        const long o = (long)(q + 1) * PF_BLK + 4 * cl;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double r = __dsub_rn(F64[o + k], U64[o + k]);
            R64[o + k] = r;
            rr[k] = __dadd_rn(rr[k], __dmul_rn(r, r));
        }
    }
    blk_block_partials(rr, partials);
}

__global__ __launch_bounds__(256) void k_blk_fold_out(const double *partials, int G, double *out)
{
    __shared__ double v[PF_BLK];
    blk_fold(partials, G, v);
    if (threadIdx.x < PF_BLK) out[threadIdx.x] = v[threadIdx.x];
}

// host order (neq x k, column-major) <-> block order ((neq+1) x 64, row 0 = 0)
__global__ void k_blk_from_host(const double *__restrict__ src, double *__restrict__ dst, int neq, int k)
{
    const long n = (long)(neq + 1) * PF_BLK;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long q = i / PF_BLK;
        const int pos = (int)(i - q * PF_BLK);
        const int c = (pos & 3) * 16 + (pos >> 2);     // inverse of pf_blk_pos
        dst[i] = (q > 0 && c < k) ? src[(long)c * neq + (q - 1)] : 0.0;
    }
}

__global__ void k_blk_to_host(const double *__restrict__ src, double *__restrict__ dst, int neq, int k)
{
    const long n = (long)neq * k;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i / neq);
        const long q = i - (long)c * neq;
        dst[i] = src[(q + 1) * PF_BLK + pf_blk_pos(c)];
    }
}

template <typename A, typename B>
__global__ void k_blk_convert(const A *__restrict__ src, B *__restrict__ dst, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = (B)src[i];
}

__global__ void k_blk_dinv(const double *__restrict__ diag, float *__restrict__ dinv, const double *__restrict__ dgiven,
                           int keep, int neq)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q <= neq; q += gridDim.x * blockDim.x)
        dinv[q] = q == 0 ? 0.0f : (float)(keep ? dgiven[q] : 1.0 / diag[q]);
}

int blk_grid(int neq)
{
    const int g = (neq + 15) / 16;
    return std::max(1, std::min(g, PF_BLK_MAXG));
}

template <typename T>
int blk_alloc(pf_ctx *c, T *&p, size_t n)
{
    return p ? PF_OK : pf_alloc(c, (void **)&p, n * sizeof(T));
}

int blk_tiles_per(int nt)
{
    switch (nt) {
    case 60: return Tiles<60>::PER;
    case 24: return Tiles<24>::PER;
    default: return Tiles<8>::PER;
    }
}

int blk_sym_per(int nt)
{
    switch (nt) {
    case 60: return SymTiles<60>::PER;
    case 24: return SymTiles<24>::PER;
    default: return SymTiles<8>::PER;
    }
}

// Allocate (once per problem) and pack the element matrices.
int blk_setup(pf_ctx *c, bool need32)
{
    if (!c->K64) return PF_ERR_ARG;            // needs FP64 element matrices
    if (!c->blk) c->blk = new pf_block();
    pf_block *b = c->blk;
    const size_t nv = (size_t)(c->neq + 1) * PF_BLK;
    const size_t ns = (size_t)c->nels * c->ntot * PF_BLK;
    const long ke = c->k_shared ? 1 : c->nels;
    const size_t nt = (size_t)ke * blk_tiles_per(c->ntot);
    const long ks = pf_k64_estride(c);
    int rc = PF_OK;
    if (!b->Kt64) {
        rc = blk_alloc(c, b->Kt64, nt);
        if (rc) return rc;
        switch (c->ntot) {
        case 60: hipLaunchKernelGGL((k_blk_pack<60, double>), dim3(4096), dim3(256), 0, c->stream, c->K64, ks, b->Kt64, ke); break;
        case 24: hipLaunchKernelGGL((k_blk_pack<24, double>), dim3(4096), dim3(256), 0, c->stream, c->K64, ks, b->Kt64, ke); break;
        default: hipLaunchKernelGGL((k_blk_pack<8, double>), dim3(4096), dim3(256), 0, c->stream, c->K64, ks, b->Kt64, ke); break;
        }
    }
    if (need32 && !b->Kt32) {
        rc = blk_alloc(c, b->Kt32, (size_t)ke * blk_sym_per(c->ntot));
        if (rc) return rc;
        switch (c->ntot) {
        case 60: hipLaunchKernelGGL(k_blk_pack_sym<60>, dim3(4096), dim3(256), 0, c->stream, c->K64, ks, b->Kt32, ke); break;
        case 24: hipLaunchKernelGGL(k_blk_pack_sym<24>, dim3(4096), dim3(256), 0, c->stream, c->K64, ks, b->Kt32, ke); break;
        default: hipLaunchKernelGGL(k_blk_pack_sym<8>, dim3(4096), dim3(256), 0, c->stream, c->K64, ks, b->Kt32, ke); break;
        }
    }
    PF_HIP(hipGetLastError());
    if (!rc) rc = blk_alloc(c, b->X64, nv);
    if (!rc) rc = blk_alloc(c, b->U64, nv);
    if (!rc) rc = blk_alloc(c, b->Ut64, ns);
    if (!rc) rc = blk_alloc(c, b->stage, (size_t)c->neq * PF_BLK);
    if (!rc) rc = blk_alloc(c, b->partials, (size_t)3 * PF_BLK_MAXG * PF_BLK);
    if (!rc && !b->st) rc = pf_alloc(c, (void **)&b->st, sizeof(pf_blk_state));
    if (!rc && !b->st_host && hipHostMalloc((void **)&b->st_host, sizeof(pf_blk_state), hipHostMallocDefault) != hipSuccess)
        rc = PF_ERR_OOM;
    if (!rc && need32) {
        rc = blk_alloc(c, b->P, nv);
        if (!rc) rc = blk_alloc(c, b->U, nv);
        if (!rc) rc = blk_alloc(c, b->R, nv);
        if (!rc) rc = blk_alloc(c, b->X, nv);
        if (!rc) rc = blk_alloc(c, b->dinv, (size_t)c->neq + 1);
        if (!rc) rc = blk_alloc(c, b->Ut32, ns);
        if (!rc) rc = blk_alloc(c, b->R64, nv);
        if (!rc) rc = blk_alloc(c, b->F64, nv);
        if (!rc) {
            // row 0 of every FP32 block vector stays zero (restrained freedoms gather it)
            PF_HIP(hipMemsetAsync(b->P, 0, PF_BLK * sizeof(float), c->stream));
            hipLaunchKernelGGL(k_blk_dinv, dim3(std::min((c->neq + 256) / 256, 8192)), dim3(256), 0, c->stream,
                               c->diag, b->dinv, c->dinv, c->precon_given ? 1 : 0, c->neq);
            PF_HIP(hipGetLastError());
        }
    }
    return rc;
}

template <int MP>
void blk_launch_mp(pf_ctx *c, int waves, const double *Kt, const double *P, double *Ut, const pf_blk_state *st)
{
    const int grid = (c->nels + waves - 1) / waves;
    const long ks = pf_k64_estride(c);
    const dim3 g(grid), b(64 * waves);
    switch (c->ntot) {
    case 60: hipLaunchKernelGGL((k_blk_ebe<60, double, MP>), g, b, 0, c->stream, Kt, ks, c->g_g, P, Ut, c->nels, st); break;
    case 24: hipLaunchKernelGGL((k_blk_ebe<24, double, MP>), g, b, 0, c->stream, Kt, ks, c->g_g, P, Ut, c->nels, st); break;
    default: hipLaunchKernelGGL((k_blk_ebe<8, double, MP>), g, b, 0, c->stream, Kt, ks, c->g_g, P, Ut, c->nels, st); break;
    }
}

// The element product in the context's launch shape: elements per workgroup
// (tune_blk_waves, default 4) and, for the FP64 product, 16-row output
// tiles per pass (tune_blk_rows, default 2); the FP32 product forms every
// row tile in one pass over the symmetric tiles.
template <typename T>
int blk_launch_ebe(pf_ctx *c, const T *Kt, const T *P, T *Ut, const pf_blk_state *st)
{
    const int waves = c->tune_blk_waves ? c->tune_blk_waves : 4;
    if constexpr (sizeof(T) == 8) {
        const int rows = c->tune_blk_rows ? c->tune_blk_rows : 2;
        if (rows == 1) blk_launch_mp<1>(c, waves, Kt, P, Ut, st);
        else if (rows == 2) blk_launch_mp<2>(c, waves, Kt, P, Ut, st);
        else blk_launch_mp<4>(c, waves, Kt, P, Ut, st);
    } else {
        const int grid = (c->nels + waves - 1) / waves;
        const long ks = pf_k64_estride(c);
        const dim3 g(grid), b(64 * waves);
        switch (c->ntot) {
        case 60: hipLaunchKernelGGL(k_blk_ebe_sym<60>, g, b, 0, c->stream, Kt, ks, c->g_g, P, Ut, c->nels, st); break;
        case 24: hipLaunchKernelGGL(k_blk_ebe_sym<24>, g, b, 0, c->stream, Kt, ks, c->g_g, P, Ut, c->nels, st); break;
        default: hipLaunchKernelGGL(k_blk_ebe_sym<8>, g, b, 0, c->stream, Kt, ks, c->g_g, P, Ut, c->nels, st); break;
        }
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
}

// U64 = K X64 (FP64 matrix cores), then R64 = F64 - U64 and per-column
// ||r||^2 into host rr[pos].
int blk_true_residual(pf_ctx *c, double *rr_host)
{
    pf_block *b = c->blk;
    const int G = blk_grid(c->neq);
    int rc = blk_launch_ebe<double>(c, b->Kt64, b->X64, b->Ut64, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL((k_blk_scatter<double, false>), dim3(G), dim3(256), 0, c->stream, c->csr_ptr, c->csr_slot,
                       b->Ut64, (const double *)nullptr, b->U64, c->neq, (double *)nullptr,
                       (const pf_blk_state *)nullptr);
    hipLaunchKernelGGL(k_blk_residual, dim3(G), dim3(256), 0, c->stream, b->R64, b->F64, b->U64, c->neq, b->partials);
    hipLaunchKernelGGL(k_blk_fold_out, dim3(1), dim3(256), 0, c->stream, b->partials, G,
                       b->partials + (long)2 * PF_BLK_MAXG * PF_BLK);
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(rr_host, b->partials + (long)2 * PF_BLK_MAXG * PF_BLK, PF_BLK * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

}  // namespace

void pf_block_free(pf_ctx *c)
{
    pf_block *b = c->blk;
    if (!b) return;
    const size_t nv = (size_t)(c->neq + 1) * PF_BLK;
    const size_t ns = (size_t)c->nels * c->ntot * PF_BLK;
    const size_t ke = (size_t)(c->k_shared ? 1 : c->nels);
    const size_t nt = ke * blk_tiles_per(c->ntot);
    (void)hipStreamSynchronize(c->stream);
    auto fr = [&](auto *&p, size_t n) {
        if (p) pf_free(c, p, n * sizeof(*p));
        p = nullptr;
    };
    fr(b->Kt32, ke * blk_sym_per(c->ntot)); fr(b->Kt64, nt);
    fr(b->P, nv); fr(b->U, nv); fr(b->R, nv); fr(b->X, nv);
    fr(b->dinv, (size_t)c->neq + 1);
    fr(b->Ut32, ns); fr(b->Ut64, ns);
    fr(b->X64, nv); fr(b->R64, nv); fr(b->F64, nv); fr(b->U64, nv);
    fr(b->stage, (size_t)c->neq * PF_BLK);
    fr(b->partials, (size_t)3 * PF_BLK_MAXG * PF_BLK);
    if (b->st) pf_free(c, b->st, sizeof(pf_blk_state));
    if (b->st_host) (void)hipHostFree(b->st_host);
    delete b;
    c->blk = nullptr;
}

extern "C" {

int pf_block_apply(pf_ctx *c, int k, const double *x, double *y, int precision)
{
    if (!c || !x || !y || k < 1 || k > PF_BLK) return PF_ERR_ARG;
    if (precision != PF_PREC_FP64 && precision != PF_PREC_MIXED) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    PF_HIP(hipSetDevice(c->device));
    const bool f32 = precision == PF_PREC_MIXED;
    int rc = blk_setup(c, f32);
    if (rc) return rc;
    pf_block *b = c->blk;
    const size_t nh = (size_t)c->neq * k;
    const int G = blk_grid(c->neq);
    PF_HIP(hipMemcpyAsync(b->stage, x, nh * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_blk_from_host, dim3(8192), dim3(256), 0, c->stream, b->stage, b->X64, c->neq, k);
    if (f32) {
        // fl32(x) -> FP32 products and scatter -> widened into U64
        hipLaunchKernelGGL((k_blk_convert<double, float>), dim3(8192), dim3(256), 0, c->stream, b->X64, b->R,
                           (long)(c->neq + 1) * PF_BLK);
        rc = blk_launch_ebe<float>(c, b->Kt32, b->R, b->Ut32, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL((k_blk_scatter<float, false>), dim3(G), dim3(256), 0, c->stream, c->csr_ptr, c->csr_slot,
                           b->Ut32, (const float *)nullptr, b->U, c->neq, (double *)nullptr,
                           (const pf_blk_state *)nullptr);
        hipLaunchKernelGGL((k_blk_convert<float, double>), dim3(8192), dim3(256), 0, c->stream, b->U, b->U64,
                           (long)(c->neq + 1) * PF_BLK);
    } else {
        rc = blk_launch_ebe<double>(c, b->Kt64, b->X64, b->Ut64, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL((k_blk_scatter<double, false>), dim3(G), dim3(256), 0, c->stream, c->csr_ptr,
                           c->csr_slot, b->Ut64, (const double *)nullptr, b->U64, c->neq, (double *)nullptr,
                           (const pf_blk_state *)nullptr);
    }
    hipLaunchKernelGGL(k_blk_to_host, dim3(8192), dim3(256), 0, c->stream, b->U64, b->stage, c->neq, k);
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(y, b->stage, nh * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_block_time_ebe(pf_ctx *c, int reps, double *ms)
{
    if (!c || !ms || reps < 1) return PF_ERR_ARG;
    if (!c->blk || !c->blk->Kt32 || !c->blk->R) return PF_ERR_STATE;
    PF_HIP(hipSetDevice(c->device));
    pf_block *b = c->blk;
    hipEvent_t t0, t1;
    PF_HIP(hipEventCreate(&t0));
    PF_HIP(hipEventCreate(&t1));
    int rc = blk_launch_ebe<float>(c, b->Kt32, b->R, b->Ut32, nullptr);   // warm
    hipError_t e = hipSuccess;
    double total = 0.0;
    // one event pair around each launch: the kernel's own duration
    for (int i = 0; !rc && e == hipSuccess && i < reps; ++i) {
        e = hipEventRecord(t0, c->stream);
        if (e == hipSuccess) rc = blk_launch_ebe<float>(c, b->Kt32, b->R, b->Ut32, nullptr);
        if (e == hipSuccess && !rc) e = hipEventRecord(t1, c->stream);
        if (e == hipSuccess && !rc) e = hipEventSynchronize(t1);
        float t = 0.0f;
        if (e == hipSuccess && !rc) e = hipEventElapsedTime(&t, t0, t1);
        total += t;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    if (e != hipSuccess) return pf_set_hip_error(e, "block timing", __FILE__, __LINE__);
    *ms = total / reps;
    return PF_OK;
}

int pf_block_solve(pf_ctx *c, int k, const double *f, double *x, const pf_block_opts *o, pf_block_result *res)
{
    if (!c || !f || !x || !o || k < 1 || k > PF_BLK) return PF_ERR_ARG;
    if (!(o->tol > 0.0) || o->inner_tol < 0.0 || o->inner_tol >= 1.0) return PF_ERR_ARG;
    if (!c->have_K) return PF_ERR_STATE;
    if (c->nfix) return PF_ERR_ARG;            // penalty rows are not part of the block operator
    PF_HIP(hipSetDevice(c->device));
    int rc = blk_setup(c, true);
    if (rc) return rc;
    pf_block *b = c->blk;
    const double delta = o->inner_tol > 0.0 ? o->inner_tol : 0.1;
    const int limit = o->inner_limit > 0 ? o->inner_limit : 100000;
    const int max_repl = o->outer_limit > 0 ? o->outer_limit : 100000;
    const int check = o->check_every > 0 ? o->check_every : 32;
    const size_t nh = (size_t)c->neq * k;
    const size_t nv = (size_t)(c->neq + 1) * PF_BLK;
    const int G = blk_grid(c->neq);
    pf_block_result R{};

    PF_HIP(hipMemcpyAsync(b->stage, f, nh * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_blk_from_host, dim3(8192), dim3(256), 0, c->stream, b->stage, b->F64, c->neq, k);
    PF_HIP(hipMemsetAsync(b->X64, 0, nv * sizeof(double), c->stream));
    PF_HIP(hipGetLastError());
    double ff[PF_BLK], rr[PF_BLK];
    rc = blk_true_residual(c, ff);             // x = 0: r = f, ||f||^2
    if (rc) return rc;
    std::copy(ff, ff + PF_BLK, rr);
    // Start: x = 0 in FP32, p = d = dinv fl32(r).
    hipLaunchKernelGGL(k_blk_inner_init<true>, dim3(G), dim3(256), 0, c->stream, b->X, b->R, b->P, b->R64, b->dinv,
                       c->neq, b->partials);
    hipLaunchKernelGGL(k_blk_fold_init, dim3(1), dim3(256), 0, c->stream, b->partials, G, b->st);
    PF_HIP(hipGetLastError());
    bool first = true;
    for (;;) {
        // Columns whose true residual meets tol (or that are unused) stop.
        int conv = 0;
        double worst = 0.0;
        PF_HIP(hipMemcpyAsync(b->st_host, b->st, sizeof(pf_blk_state), hipMemcpyDeviceToHost, c->stream));
        PF_HIP(hipStreamSynchronize(c->stream));
        pf_blk_state s0 = *b->st_host;
        for (int col = 0; col < PF_BLK; ++col) {
            const int p = pf_blk_pos(col);
            const double rel = ff[p] > 0.0 ? std::sqrt(rr[p] / ff[p]) : 0.0;
            const bool done = col >= k || rel <= o->tol;
            if (col < k) {
                worst = std::max(worst, rel);
                conv += done;
            }
            s0.frozen[p] = done;
            s0.ff[p] = ff[p];
        }
        R.converged = conv;
        R.max_relres = worst;
        if (conv == k || R.outer >= max_repl || (!first && s0.iters >= limit)) break;
        if (first) {
            s0.iters = 0;
            s0.limit = limit;
            s0.done = 0;
        }
        s0.tol2 = o->tol * o->tol;
        s0.delta2 = delta * delta;
        s0.trig = 0;
        s0.xlast = 0;
        *b->st_host = s0;
        PF_HIP(hipMemcpyAsync(b->st, b->st_host, sizeof(pf_blk_state), hipMemcpyHostToDevice, c->stream));
        first = false;
        // FP32 block PCG until some column asks for a residual replacement.
        for (;;) {
            for (int it = 0; it < check; ++it) {
                // with timing on, one event pair per kernel group (pf_pcg_timing_get)
                hipEvent_t t0, t1;
                bool tm = pf_timing_begin(c, PF_K_EBE, &t0, &t1);
                if (tm) PF_HIP(hipEventRecord(t0, c->stream));
                rc = blk_launch_ebe<float>(c, b->Kt32, b->P, b->Ut32, b->st);
                if (rc) return rc;
                if (tm) PF_HIP(hipEventRecord(t1, c->stream));
                tm = pf_timing_begin(c, PF_K_SCATTER, &t0, &t1);
                if (tm) PF_HIP(hipEventRecord(t0, c->stream));
                hipLaunchKernelGGL((k_blk_scatter<float, true>), dim3(G), dim3(256), 0, c->stream, c->csr_ptr,
                                   c->csr_slot, b->Ut32, b->P, b->U, c->neq, b->partials, b->st);
                hipLaunchKernelGGL(k_blk_fold_alpha, dim3(1), dim3(256), 0, c->stream, b->partials, G, b->st);
                if (tm) PF_HIP(hipEventRecord(t1, c->stream));
                tm = pf_timing_begin(c, PF_K_UPDATE, &t0, &t1);
                if (tm) PF_HIP(hipEventRecord(t0, c->stream));
                hipLaunchKernelGGL(k_blk_update, dim3(G), dim3(256), 0, c->stream, b->R, b->U, b->dinv, c->neq,
                                   b->partials, b->st);
                hipLaunchKernelGGL(k_blk_fold_update, dim3(1), dim3(256), 0, c->stream, b->partials, G, b->st);
                if (tm) PF_HIP(hipEventRecord(t1, c->stream));
                tm = pf_timing_begin(c, PF_K_PUPDATE, &t0, &t1);
                if (tm) PF_HIP(hipEventRecord(t0, c->stream));
                hipLaunchKernelGGL(k_blk_pupdate, dim3(G), dim3(256), 0, c->stream, b->X, b->P, b->R, b->dinv,
                                   c->neq, b->st);
                if (tm) PF_HIP(hipEventRecord(t1, c->stream));
            }
            PF_HIP(hipGetLastError());
            PF_HIP(hipMemcpyAsync(b->st_host, b->st, sizeof(pf_blk_state), hipMemcpyDeviceToHost, c->stream));
            PF_HIP(hipStreamSynchronize(c->stream));
            if (b->st_host->trig || b->st_host->done) break;
        }
        R.inner_iters = b->st_host->iters;
        // Residual replacement: x += the FP32 correction, r = f - K x in FP64,
        // the FP32 recurrence restarts from fl32(r) with the same p.
        hipLaunchKernelGGL(k_blk_correct<true>, dim3(8192), dim3(256), 0, c->stream, b->X64, b->X,
                           (long)c->neq * PF_BLK);
        PF_HIP(hipGetLastError());
        double prev[PF_BLK];
        std::copy(rr, rr + PF_BLK, prev);
        rc = blk_true_residual(c, rr);
        if (rc) return rc;
        // A column whose true residual grew since the last replacement has
        // search directions spoilt by FP32 rounding: every column restarts
        // from p = dinv r (x is already in the FP64 sum).
        bool grew = false;
        for (int col = 0; col < k; ++col) {
            const int p = pf_blk_pos(col);
            grew |= !b->st_host->frozen[p] && rr[p] > 4.0 * prev[p];
        }
        if (grew)
            hipLaunchKernelGGL(k_blk_inner_init<true>, dim3(G), dim3(256), 0, c->stream, b->X, b->R, b->P, b->R64,
                               b->dinv, c->neq, b->partials);
        else
            hipLaunchKernelGGL(k_blk_inner_init<false>, dim3(G), dim3(256), 0, c->stream, b->X, b->R, b->P, b->R64,
                               b->dinv, c->neq, b->partials);
        hipLaunchKernelGGL(k_blk_fold_init, dim3(1), dim3(256), 0, c->stream, b->partials, G, b->st);
        PF_HIP(hipGetLastError());
        R.outer += 1;
    }
    hipLaunchKernelGGL(k_blk_to_host, dim3(8192), dim3(256), 0, c->stream, b->X64, b->stage, c->neq, k);
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(x, b->stage, nh * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(hipStreamSynchronize(c->stream));
    R.status = R.converged == k ? PF_OK : PF_NOT_CONVERGED;
    if (res) *res = R;
    return R.status;
}

}  // extern "C"
