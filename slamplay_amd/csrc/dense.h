// dense.h -- explicit reduced camera system and its dense Cholesky solve (the DENSE_SCHUR /
// SPARSE_SCHUR step of test_bundle_adjustment_ceres.cpp:97 for problems up to a few hundred cameras).
//
//   S = U* - sum_points sum_{i,j in track} W~_i Vinv_p W~_j'     (9 nc x 9 nc, FP64)
// with U*'s per-observation part folded into the (i, i) pairs: S_aa = D_a^2 + sum J~'(I - G)J~ - ...
// is assembled without atomics: the host enumerates every ordered observation pair (i, j) of a
// point with camera(i) <= camera(j) once per problem and groups the pairs by camera block (a, b);
// one wave owns one 9x9 block and reduces its pair list in a fixed order (pair_block.h, the kernel
// the batched windows use too).  A pair contributes Jc~_i' (Hp_i Jp~_j') Jc~_j with
// Hp_i = Jp~_i Vinv_p, i.e. a 2x2 middle factor.
//
// The factorization is a tiled right-looking Cholesky (32 x 32 tiles): per step, the diagonal
// tile is factored by one wave and its triangular inverse kept, the panel below is multiplied by
// that inverse (TRSM as a small GEMM) and the trailing lower tiles are updated (SYRK/GEMM); both
// products run on the matrix cores (v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32), and the
// trailing update also factors the next diagonal tile, so a step is two launches.  The panel is
// also stored transposed in the upper tiles (the factorization reads only the lower ones; in FP64
// the factor overwrites S).  The triangular solves work on diagonal blocks of four tiles whose
// inverses are formed after the factorization, one launch per block and sweep, with the rows
// below (above) the block updated by many workgroups at once.  In mixed precision the factor is
// FP32 and the step is refined in FP64 against the FP64 S (refine_steps iterations of
// r = b - S y, y += S_f^-1 r).
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "lm_kernels.h"
#include "mfma.h"
#include "pair_block.h"
#include "solver.h"

namespace mba {

constexpr int kTile = 32;
constexpr int kJs = 36;  // per-observation stride of Jc~ (18, by columns), Jp~ (6), Hp (6), I - G (3)
static_assert(kJs == RecLayout<kRecDense>::stride, "the dense record is pair_block's kRecDense");

template <typename Real, int M>
__global__ void __launch_bounds__(kWG) k_dense_store(Dev<Real> d, Real* __restrict__ js) {
    LMState* st = d.st;
    if (st->done || st->step_invalid) return;
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    const CameraPrep<Real> P = d.prep[c];
    Real sc[9];
    for (int i = 0; i < 9; ++i) sc[i] = (Real)d.sc[9 * (size_t)c + i];
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6];
        corrected_jacobian<Real, M>(d, P, j, p, r, Jc, Jp);
        Real* o = js + kJs * (size_t)j;
        for (int i = 0; i < 9; ++i) {  // J~_c by columns: the two residual rows of column i adjacent
            o[2 * i] = Jc[i] * sc[i];
            o[2 * i + 1] = Jc[9 + i] * sc[i];
        }
        for (int i = 0; i < 3; ++i) {
            const Real s = (Real)d.sp[3 * (size_t)p + i];
            Jp[i] *= s;
            Jp[3 + i] *= s;
        }
        for (int i = 0; i < 6; ++i) o[18 + i] = Jp[i];
        Real h[6];
        sym3_mul<Real>(d.VinvR + 6 * (size_t)p, Jp, h);
        sym3_mul<Real>(d.VinvR + 6 * (size_t)p, Jp + 3, h + 3);
        for (int i = 0; i < 6; ++i) o[24 + i] = h[i];
        point_complement<Real, Real>(d, p, Jp, o[30], o[31], o[32]);
    }
}

// S <- 0 with an identity tail on the padding diagonal.
__global__ void k_fill_S(double* __restrict__ S, int n, int ld, const LMState* st) {
    if (st->done || st->step_invalid) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ld * ld) return;
    const int r = (int)(i / ld), c = (int)(i % ld);
    S[i] = (r == c && r >= n) ? 1.0 : 0.0;
}

// One 9x9 block of the reduced camera system per workgroup of one wave (the blocks' pair lists, the
// 36-value records of k_dense_store): pair_block in pair_block.h.
template <typename Real>
__global__ void __launch_bounds__(64) k_schur_blocks(Dev<Real> d, const Real* __restrict__ js,
                                                     const int* __restrict__ blk_a, const int* __restrict__ blk_b,
                                                     const int* __restrict__ blk_off, const int2* __restrict__ pairs,
                                                     double* __restrict__ S, int ld, int nblocks) {
    LMState* st = d.st;
    if (st->done || st->step_invalid) return;
    __shared__ __align__(16) unsigned char stage[pair_stage_bytes<Real, kRecDense>()];
    const int b = blockIdx.x;
    if (b >= nblocks) return;
    const int4 bk = make_int4(blk_a[b], blk_b[b], blk_off[b], blk_off[b + 1]);
    pair_block<Real, kRecDense>(bk, pairs, js, d.dc2, S, ld, reinterpret_cast<Real*>(stage));
}

template <typename F>
__global__ void k_convert_S(const double* __restrict__ S, F* __restrict__ A, size_t n, const LMState* st) {
    if (st->done || st->step_invalid) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[i] = (F)S[i];
}

// ---------------------------------------------------------------------------------------------
// 32 x 32 tile products on the matrix cores.  A 16 x 16 block of the product X Y' of two 32-column
// tiles is one wave's eight v_mfma 16x16x4 steps: in step s, lane l supplies X[row0 + l%16][8q + s]
// and Y[col0 + l%16][8q + s] (q = l/16), so each lane reads 8 consecutive values of one row and the
// four lanes of a row cover it; every k in [0, 32) enters exactly once.  Accumulator element i of
// lane l is entry (row(l, i), l%16) of the block; the FP64 and FP32 instructions order rows
// differently.

// c + sign * X[16 si .., :] Y[16 sj .., :]' for 32-column tiles X (leading dimension ldx) and Y (ldy);
// rows start 8-element aligned.
template <typename F>
__device__ __forceinline__ typename Mfma<F>::acc tile_mma_nt(const F* __restrict__ X, int ldx, const F* __restrict__ Y,
                                                             int ldy, int si, int sj, typename Mfma<F>::acc c, F sign) {
    const int l = threadIdx.x & 63, q = l >> 4, r = l & 15;
    F xv[8], yv[8];
    Mfma<F>::load8(X + (size_t)(16 * si + r) * ldx + 8 * q, xv);
    Mfma<F>::load8(Y + (size_t)(16 * sj + r) * ldy + 8 * q, yv);
#pragma unroll
    for (int s = 0; s < 8; ++s) c = Mfma<F>::mma(sign * xv[s], yv[s], c);
    return c;
}

// Factors the tile held in T (lower triangle used) with one wave and writes L (lower triangle of
// diagonal tile k of A) and X = L^-1 (Dinv[k], row-major, zero above the diagonal).  Lane r < 32
// keeps row r in registers; the pivot and the entries l_cj of the scaled column j are the same
// for every lane and are read from their lanes with v_readlane, so the factorization makes no
// trip through LDS.  X is formed by columns: lane c < 32 solves L x = e_c by forward substitution
// (x[m] = 0 for m < c follows by itself) with L broadcast from LDS.  Returns false (and writes
// nothing) when a pivot is not positive and finite.  Only wave 0 may call it; lanes 32..63 repeat
// the work of lanes 0..31 and store nothing.
template <typename F>
__device__ __forceinline__ bool potrf_tile_wave(F (*T)[kTile + 1], F* dl, F* __restrict__ A, int ld, int k,
                                                F* __restrict__ Dinv) {
    const int l = threadIdx.x & 63, r = l & 31;
    F a[kTile];
#pragma unroll
    for (int c = 0; c < kTile; ++c) a[c] = T[r][c];
    bool bad = false;
#pragma unroll
    for (int j = 0; j < kTile; ++j) {
        const F d = read_lane(a[j], j);
        const bool nonpos = !(d > (F)0) || !isfinite((double)d);
        bad |= nonpos;
        const F dd = nonpos ? (F)1 : d, inv = inv_sqrt(dd), ljj = dd * inv;
        a[j] = r > j ? a[j] * inv : (r == j ? ljj : (F)0);
        if (l == 0) dl[j] = inv;
#pragma unroll
        for (int c = j + 1; c < kTile; ++c) {
            const F lcj = read_lane(a[j], c);
            if (r >= c) a[c] -= a[j] * lcj;
        }
    }
    if (bad) return false;
#pragma unroll
    for (int c = 0; c < kTile; ++c) T[r][c] = c <= r ? a[c] : (F)0;
    wave_sync_lds();
    const size_t o = (size_t)(kTile * k) * ld + kTile * k;
    for (int idx = l; idx < kTile * kTile; idx += 64) {
        const int rr = idx / kTile, cc = idx % kTile;
        if (cc <= rr) A[o + (size_t)rr * ld + cc] = T[rr][cc];
    }
    F x[kTile];
#pragma unroll
    for (int m = 0; m < kTile; ++m) {
        F p0 = 0, p1 = 0;
#pragma unroll
        for (int q = 0; q + 1 < m; q += 2) {
            p0 += T[m][q] * x[q];
            p1 += T[m][q + 1] * x[q + 1];
        }
        if (m & 1) p0 += T[m][m - 1] * x[m - 1];
        x[m] = ((m == r) ? (F)1 - (p0 + p1) : -(p0 + p1)) * dl[m];
    }
    if (l < kTile)
#pragma unroll
        for (int m = 0; m < kTile; ++m) Dinv[(size_t)k * kTile * kTile + m * kTile + r] = x[m];
    return true;
}

// Diagonal tile k (used for the first tile; later ones are factored by the trailing update).
template <typename F>
__global__ void __launch_bounds__(64) k_potrf_tile(F* __restrict__ A, int ld, int k, F* __restrict__ Dinv, LMState* st) {
    if (st->done || st->step_invalid) return;
    __shared__ F T[kTile][kTile + 1];
    __shared__ F dl[kTile];
    const int l = threadIdx.x;
    const size_t o = (size_t)(kTile * k) * ld + kTile * k;
    for (int idx = l; idx < kTile * kTile; idx += 64) T[idx / kTile][idx % kTile] = A[o + (size_t)(idx / kTile) * ld + idx % kTile];
    __syncthreads();
    if (!potrf_tile_wave<F>(T, dl, A, ld, k, Dinv) && l == 0) st->step_invalid = 1;
}

// Step k of the factorization in one launch, one workgroup per lower tile (i, j) of the trailing
// matrix, k < j <= i, one 16 x 16 block per wave.  The panel tiles the update needs are formed in
// the workgroup from the unscaled A_ik, A_jk and the kept inverse X_k = L_kk^-1 (L_ik = A_ik X_k',
// the same matrix-core products in the same order as a separate panel launch would do), staged in
// LDS, and A_ij -= L_ik L_jk' follows; so a step is one launch.  The workgroups of column k + 1
// store L_ik' in the upper tile (k, i), which the factorization does not read; the lower tile
// (i, k) still holds A_ik, which other workgroups of this launch read, and k_panel_down copies the
// transposes into it once the factorization is done.  Block 0 (tile (k + 1, k + 1), dispatched
// first) keeps its result in LDS and factors it, so the next step's diagonal factor costs no launch
// of its own.
template <typename F>
__global__ void __launch_bounds__(256) k_step_tiles(F* __restrict__ A, int ld, int k, F* __restrict__ Dinv, LMState* st) {
    if (st->done || st->step_invalid) return;
    constexpr int PP = kTile + 16 / (int)sizeof(F);  // LDS row pitch: rows start 16-byte aligned
    __shared__ __align__(16) F P[2][kTile * PP];     // L_ik, L_jk
    __shared__ F T[kTile][kTile + 1];
    __shared__ F dl[kTile];
    const int b = blockIdx.x, w = threadIdx.x >> 6, l = threadIdx.x & 63;
    int ii = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((ii + 1) * (ii + 2) / 2 <= b) ++ii;
    while (ii * (ii + 1) / 2 > b) --ii;
    const int jj = b - ii * (ii + 1) / 2;
    const int i = k + 1 + ii, j = k + 1 + jj;
    const int si = w >> 1, sj = w & 1;
    const bool next_diag = b == 0;
    const F* Xk = Dinv + (size_t)k * kTile * kTile;
    // panel tiles of rows i and j (one of them when i == j)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        if (t == 1 && i == j) break;
        const F* Ark = A + (size_t)(kTile * (t == 0 ? i : j)) * ld + kTile * k;
        typename Mfma<F>::acc c = {0, 0, 0, 0};
        c = tile_mma_nt<F>(Ark, ld, Xk, kTile, si, sj, c, (F)1);
#pragma unroll
        for (int e = 0; e < 4; ++e) P[t][(16 * si + Mfma<F>::row(l, e)) * PP + 16 * sj + (l & 15)] = c[e];
    }
    __syncthreads();
    if (j == k + 1) {
        F* Aki = A + (size_t)(kTile * k) * ld + kTile * i;
        for (int idx = threadIdx.x; idx < kTile * kTile; idx += 256) {
            const int cc = idx / kTile, rr = idx % kTile;
            Aki[(size_t)cc * ld + rr] = P[0][rr * PP + cc];
        }
    }
    F* Aij = A + (size_t)(kTile * i) * ld + kTile * j;
    if (!(i == j && si < sj)) {
        typename Mfma<F>::acc c;
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = Aij[(size_t)(16 * si + Mfma<F>::row(l, e)) * ld + 16 * sj + (l & 15)];
        c = tile_mma_nt<F>(P[0], PP, i == j ? P[0] : P[1], PP, si, sj, c, (F)-1);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int rr = 16 * si + Mfma<F>::row(l, e), cc = 16 * sj + (l & 15);
            if (next_diag)
                T[rr][cc] = c[e];
            else
                Aij[(size_t)rr * ld + cc] = c[e];
        }
    }
    if (!next_diag) return;
    __syncthreads();
    if (w == 0 && !potrf_tile_wave<F>(T, dl, A, ld, k + 1, Dinv) && l == 0) st->step_invalid = 1;
}

// After the factorization: L_ik = (tile (k, i))' into the lower tile (i, k), i > k (one workgroup per
// tile; the triangular sweeps and the block inverses read the lower tiles).
template <typename F>
__global__ void __launch_bounds__(256) k_panel_down(F* __restrict__ A, int ld, const LMState* st) {
    if (st->done || st->step_invalid) return;
    __shared__ F T[kTile][kTile + 1];
    const int b = blockIdx.x;
    int ii = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((ii + 1) * (ii + 2) / 2 <= b) ++ii;
    while (ii * (ii + 1) / 2 > b) --ii;
    const int k = b - ii * (ii + 1) / 2, i = ii + 1;  // k < i
    const F* Aki = A + (size_t)(kTile * k) * ld + kTile * i;
    F* Aik = A + (size_t)(kTile * i) * ld + kTile * k;
    for (int idx = threadIdx.x; idx < kTile * kTile; idx += 256) T[idx / kTile][idx % kTile] = Aki[(size_t)(idx / kTile) * ld + idx % kTile];
    __syncthreads();
    for (int idx = threadIdx.x; idx < kTile * kTile; idx += 256) Aik[(size_t)(idx / kTile) * ld + idx % kTile] = T[idx % kTile][idx / kTile];
}

// ---------------------------------------------------------------------------------------------
// Blocked triangular solves over many workgroups.  The factor is cut into diagonal blocks of
// kSolveBlk tiles (128 rows); after the factorization k_block_inverse forms each block's inverse
// X = L_KK^-1 from the kept tile inverses (X_ii = Dinv_i, X_ij = -Dinv_i sum_{j<=k<i} L_ik X_kj),
// so a block of either sweep is one matrix-vector product.  Forward sweep, launch K: workgroup 0
// applies block K-1's solution to block K's rows and solves block K (y_K = X_K z_K); the other
// workgroups apply block K-1's solution to one tile row each below block K.  The backward sweep
// runs the same way upwards with the transposed panel (upper tiles) and x_K = X_K' z_K.  One sweep
// is one launch per block; every workgroup streams rows of the factor with one wave per row.
constexpr int kSolveBlk = 4;
constexpr int kSolveRows = kSolveBlk * kTile;

// Tiles of block K: [kSolveBlk K, min(kSolveBlk (K + 1), nt)).
__device__ __forceinline__ int solve_block_tiles(int K, int nt) { return min(kSolveBlk, nt - kSolveBlk * K); }

// One workgroup per (block K, column tile j): column j of X = L_KK^-1, i.e. X_jj = Dinv_j and
// X_ij = -Dinv_i sum_{j<=k<i} L_ik X_kj for i = j+1, ... in order, with the column's tiles, the
// current factor tile and the partial sum staged in LDS (FP64).  X is stored row-major and, for the
// backward sweep, transposed (XT), so both sweeps read rows.
template <typename F>
__global__ void __launch_bounds__(256) k_block_inverse(const F* __restrict__ A, int ld, const F* __restrict__ Dinv,
                                                        F* __restrict__ Binv, F* __restrict__ BinvT, int nt,
                                                        const LMState* st) {
    if (st->done || st->step_invalid) return;
    __shared__ double Xc[kSolveBlk][kTile][kTile + 1];
    __shared__ double Lt[kTile][kTile + 1];
    __shared__ double Tt[kTile][kTile + 1];
    const int K = blockIdx.x / kSolveBlk, j = blockIdx.x % kSolveBlk;
    const int bt = solve_block_tiles(K, nt), t0 = kSolveBlk * K;
    const int tid = threadIdx.x;
    F* X = Binv + (size_t)K * kSolveRows * kSolveRows;
    F* XT = BinvT + (size_t)K * kSolveRows * kSolveRows;
    // this workgroup's column: zero above the diagonal tile (and the whole column past the block)
    for (int e = tid; e < kSolveRows * kTile; e += 256) {
        const int r = e / kTile, c = kTile * j + e % kTile;
        if (r < kTile * j || j >= bt || r >= kTile * bt) {
            X[(size_t)r * kSolveRows + c] = (F)0;
            XT[(size_t)c * kSolveRows + r] = (F)0;
        }
    }
    if (j >= bt) return;
    auto put = [&](int i, int r, int c, double v) {
        Xc[i][r][c] = v;
        X[(size_t)(kTile * i + r) * kSolveRows + kTile * j + c] = (F)v;
        XT[(size_t)(kTile * j + c) * kSolveRows + kTile * i + r] = (F)v;
    };
    for (int e = tid; e < kTile * kTile; e += 256)
        put(j, e / kTile, e % kTile, (double)Dinv[(size_t)(t0 + j) * kTile * kTile + e]);
    __syncthreads();
    for (int i = j + 1; i < bt; ++i) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = j; k < i; ++k) {
            const F* Lik = A + (size_t)(kTile * (t0 + i)) * ld + kTile * (t0 + k);
            for (int e = tid; e < kTile * kTile; e += 256) Lt[e / kTile][e % kTile] = (double)Lik[(size_t)(e / kTile) * ld + e % kTile];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = tid + 256 * q, r = e / kTile, c = e % kTile;
                double sum = 0.0;
#pragma unroll 8
                for (int m = 0; m < kTile; ++m) sum += Lt[r][m] * Xc[k][m][c];
                acc[q] += sum;
            }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            Tt[e / kTile][e % kTile] = acc[q];
        }
        const F* Di = Dinv + (size_t)(t0 + i) * kTile * kTile;
        for (int e = tid; e < kTile * kTile; e += 256) Lt[e / kTile][e % kTile] = (double)Di[e];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q, r = e / kTile, c = e % kTile;
            double sum = 0.0;
#pragma unroll 8
            for (int m = 0; m < kTile; ++m) sum += Lt[r][m] * Tt[m][c];
            put(i, r, c, -sum);
        }
        __syncthreads();
    }
}

// Row dot products out(rr, dot(M[r0 + rr][c0 .. c0 + ncols), v)) for rr < nrows of a row-major M
// (leading dimension ldm), ncols <= 128, v in double.  One wave per row: lane l takes columns l and
// l + 64, the wave sums; a wave takes kRowBatch rows at a time with all their loads in flight
// before the first sum, so the rows do not wait on each other's loads.
constexpr int kRowBatch = 16;

template <typename F, typename Out>
__device__ __forceinline__ void rows_dot(const F* __restrict__ M, size_t ldm, int r0, int nrows, int c0, int ncols,
                                         const double* __restrict__ v, Out out) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double v0 = lane < ncols ? v[lane] : 0.0, v1 = lane + 64 < ncols ? v[lane + 64] : 0.0;
    // column and row indices are clamped into the block instead of branching around the loads (a
    // branch per load makes every load wait for its own result); clamped columns carry v = 0 and
    // clamped rows are not written
    const int ca = min(lane, ncols - 1), cb = min(lane + 64, ncols - 1);
    const int nw = blockDim.x >> 6;
    for (int rb = wave * kRowBatch; rb < nrows; rb += nw * kRowBatch) {
        F x0[kRowBatch], x1[kRowBatch];
#pragma unroll
        for (int g = 0; g < kRowBatch; ++g) {
            const F* row = M + (size_t)(r0 + min(rb + g, nrows - 1)) * ldm + c0;
            x0[g] = row[ca];
            x1[g] = row[cb];
        }
        double s[kRowBatch];
#pragma unroll
        for (int g = 0; g < kRowBatch; ++g) {
            double a = (double)x0[g] * v0;
            a += (double)x1[g] * v1;
            s[g] = a;
        }
#pragma unroll
        for (int g = 0; g < kRowBatch; ++g) s[g] = wave_sum(s[g]);
        if (lane == 0)
#pragma unroll
            for (int g = 0; g < kRowBatch; ++g)
                if (rb + g < nrows) out(rb + g, s[g]);
    }
}

// z[r0 + rr] -= M[r0 + rr][c0 ..] v for rr < nrows.
template <typename F>
__device__ __forceinline__ void rows_minus_mv(const F* __restrict__ M, size_t ldm, int r0, int nrows, int c0,
                                              int ncols, const double* __restrict__ v, double* __restrict__ z) {
    rows_dot<F>(M, ldm, r0, nrows, c0, ncols, v, [&](int rr, double s) { z[r0 + rr] -= s; });
}

// Forward sweep, launch K (see above).  z has ld entries; launch 0 copies b into z first.  The
// diagonal workgroup keeps block K's right-hand side in LDS between its two products.
constexpr int kSolveThreads = 512;

template <typename F>
__global__ void __launch_bounds__(kSolveThreads) k_trsv_fwd(const F* __restrict__ A, int ld, const F* __restrict__ Binv,
                                                            int nt, int K, const double* __restrict__ b, int n,
                                                            double* __restrict__ z, const LMState* st) {
    if (st->done || st->step_invalid) return;
    __shared__ double v[kSolveRows];
    __shared__ double zk[kSolveRows];
    const int t = threadIdx.x, g = blockIdx.x;
    const int kb = kSolveRows * K;                       // first row of block K
    const int rows_k = kTile * solve_block_tiles(K, nt);  // rows of block K
    const int pb = kb - kSolveRows;                      // block K-1 (always full when K > 0)
    if (g > 0) {
        // one tile row below block K
        const int r0 = kb + rows_k + kTile * (g - 1);
        if (K == 0) {
            for (int r = t; r < kTile; r += blockDim.x) z[r0 + r] = r0 + r < n ? b[r0 + r] : 0.0;
            return;
        }
        for (int c = t; c < kSolveRows; c += blockDim.x) v[c] = z[pb + c];
        __syncthreads();
        rows_minus_mv<F>(A, (size_t)ld, r0, kTile, pb, kSolveRows, v, z);
        return;
    }
    for (int r = t; r < rows_k; r += blockDim.x) zk[r] = K == 0 ? (r < n ? b[r] : 0.0) : z[kb + r];
    if (K > 0)
        for (int c = t; c < kSolveRows; c += blockDim.x) v[c] = z[pb + c];
    __syncthreads();
    if (K > 0) {
        rows_dot<F>(A, (size_t)ld, kb, rows_k, pb, kSolveRows, v, [&](int rr, double s) { zk[rr] -= s; });
        __syncthreads();
    }
    // y_K = X_K z_K
    const F* X = Binv + (size_t)K * kSolveRows * kSolveRows;
    rows_dot<F>(X, (size_t)kSolveRows, 0, rows_k, 0, rows_k, zk, [&](int r, double s) { z[kb + r] = s; });
}

// Backward sweep, launch K (blocks from the last down to 0): x_K = X_K' (z_K - U_{K,K+1} x_{K+1});
// rows above block K get block K+1's solution applied.  Block K's result also goes to x (added to
// it when accumulate != 0), rows below n excluded.
template <typename F>
__global__ void __launch_bounds__(kSolveThreads) k_trsv_bwd(const F* __restrict__ A, int ld, const F* __restrict__ BinvT,
                                                            int nt, int K, double* __restrict__ z, double* __restrict__ x,
                                                            int n, int accumulate, const LMState* st) {
    if (st->done || st->step_invalid) return;
    __shared__ double v[kSolveRows];
    __shared__ double zk[kSolveRows];
    const int t = threadIdx.x, g = blockIdx.x;
    const int nb = (nt + kSolveBlk - 1) / kSolveBlk;
    const int kb = kSolveRows * K, rows_k = kTile * solve_block_tiles(K, nt);
    const bool has_next = K + 1 < nb;
    const int nbk = has_next ? kSolveRows * (K + 1) : 0;                      // first row of block K+1
    const int cols_next = has_next ? kTile * solve_block_tiles(K + 1, nt) : 0;  // its rows
    if (g > 0) {
        if (!has_next) return;
        const int r0 = kTile * (g - 1);  // a tile row above block K
        for (int c = t; c < cols_next; c += blockDim.x) v[c] = z[nbk + c];
        __syncthreads();
        rows_minus_mv<F>(A, (size_t)ld, r0, kTile, nbk, cols_next, v, z);
        return;
    }
    for (int r = t; r < rows_k; r += blockDim.x) zk[r] = z[kb + r];
    if (has_next)
        for (int c = t; c < cols_next; c += blockDim.x) v[c] = z[nbk + c];
    __syncthreads();
    if (has_next) {
        rows_dot<F>(A, (size_t)ld, kb, rows_k, nbk, cols_next, v, [&](int rr, double s) { zk[rr] -= s; });
        __syncthreads();
    }
    // x_K = X_K' z_K (rows of the transposed inverse)
    const F* XT = BinvT + (size_t)K * kSolveRows * kSolveRows;
    rows_dot<F>(XT, (size_t)kSolveRows, 0, rows_k, 0, rows_k, zk, [&](int r, double s) {
        z[kb + r] = s;
        const int row = kb + r;
        if (row < n) x[row] = accumulate ? x[row] + s : s;
    });
}

// r = b - S y (FP64, S symmetric with its full n x n block stored); one wave per row.
__global__ void __launch_bounds__(kWG) k_residual(const double* __restrict__ S, int ld, const double* __restrict__ b,
                                                  const double* __restrict__ y, double* __restrict__ r, int n,
                                                  const LMState* st) {
    if (st->done || st->step_invalid) return;
    const int row = blockIdx.x * kWaves + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    double v = 0.0;
    for (int c = lane; c < n; c += 64) v += S[(size_t)row * ld + c] * y[c];
    v = wave_sum(v);
    if (lane == 0) r[row] = b[row] - v;
}

__global__ void k_set_linear_iters(LMState* st, int iters) {
    if (st->done || st->step_invalid) return;
    st->pcg_iters = iters;
}

// Reduced systems up to this size are factored in FP64 in both precisions: the factor costs a few
// microseconds there, and an FP32 factor of a small, nearly gauge-free system (a local window whose
// trust region has grown) would need a Krylov solve on top to reach the FP64 step.
constexpr int kDenseF64MaxN = 288;

template <typename Real>
struct DenseSchur {
    int n = 0, ld = 0, nt = 0, nb = 0, nblocks = 0, refine = 0;
    bool f64 = true;  // factor in FP64 (FP64 solver, or a small system)
    DevBuf<Real> js;
    DevBuf<int> blk_a, blk_b, blk_off;
    DevBuf<int2> pairs;
    DevBuf<double> S, r, z;
    DevBuf<float> Af, Dinvf, Binvf, BinvTf;
    DevBuf<double> Dinvd, Binvd, BinvTd;
    int64_t device_bytes() const {
        return js.bytes() + blk_a.bytes() * 3 + pairs.bytes() + S.bytes() + r.bytes() + z.bytes() + Af.bytes() +
               Dinvf.bytes() + Binvf.bytes() + BinvTf.bytes() + Dinvd.bytes() + Binvd.bytes() + BinvTd.bytes();
    }

    int create(const ba_problem_view* v, const Layout& L, const ba_options& o, hipStream_t s) {
        (void)v;
        const int nc = L.nc;
        n = 9 * nc;
        nt = std::max(1, (n + kTile - 1) / kTile);
        ld = nt * kTile;
        nb = (nt + kSolveBlk - 1) / kSolveBlk;
        f64 = sizeof(Real) == sizeof(double) || n <= kDenseF64MaxN;
        refine = f64 ? 0 : std::max(0, o.refine_steps);
        PairBlocks pb;
        build_pair_blocks(L, &pb);
        const std::vector<int>&ba = pb.a, &bb = pb.b, &boff = pb.off, &pi = pb.i, &pj = pb.j;
        nblocks = (int)ba.size();
        BA_TRY(blk_a.upload(ba.data(), ba.size(), s));
        BA_TRY(blk_b.upload(bb.data(), bb.size(), s));
        BA_TRY(blk_off.upload(boff.data(), boff.size(), s));
        std::vector<int2> pr(pi.size());
        for (size_t q = 0; q < pi.size(); ++q) pr[q] = make_int2(pi[q], pj[q]);
        BA_TRY(pairs.upload(pr.data(), pr.size(), s));
        BA_TRY(js.alloc(kJs * (size_t)std::max(L.no, 1)));
        BA_TRY(S.alloc((size_t)ld * ld));
        BA_TRY(r.alloc(ld));
        BA_TRY(z.alloc(ld));
        const size_t binv = (size_t)nb * kSolveRows * kSolveRows;
        if (f64) {
            BA_TRY(Dinvd.alloc((size_t)nt * kTile * kTile));
            BA_TRY(Binvd.alloc(binv));
            BA_TRY(BinvTd.alloc(binv));
        } else {
            BA_TRY(Af.alloc((size_t)ld * ld));
            BA_TRY(Dinvf.alloc((size_t)nt * kTile * kTile));
            BA_TRY(Binvf.alloc(binv));
            BA_TRY(BinvTf.alloc(binv));
        }
        BA_HIP(hipStreamSynchronize(s));
        return BA_OK;
    }

    // x = (L L')^-1 b (accumulate: x += ...) by the blocked sweeps: nb launches each way.
    template <typename F, typename SolverT>
    void solve(SolverT& sv, const F* A, const F* Binv, const F* BinvT, const double* b, double* x, int accumulate) {
        auto& d = sv.d;
        for (int K = 0; K < nb; ++K) {
            const int below = std::max(0, nt - kSolveBlk * (K + 1));
            sv.launch("trsv_fwd", k_trsv_fwd<F>, 1 + below, kSolveThreads, A, ld, Binv, nt, K, b, n, z.p,
                      (const LMState*)d.st);
        }
        for (int K = nb - 1; K >= 0; --K)
            sv.launch("trsv_bwd", k_trsv_bwd<F>, 1 + kSolveBlk * K, kSolveThreads, A, ld, BinvT, nt, K, z.p, x, n,
                      accumulate, (const LMState*)d.st);
    }

    template <typename F, typename SolverT>
    void factor_and_solve(SolverT& sv, F* A, F* Dinv, F* Binv, F* BinvT) {
        auto& d = sv.d;
        sv.launch("potrf", k_potrf_tile<F>, 1, 64, A, ld, 0, Dinv, d.st);
        for (int k = 0; k + 1 < nt; ++k) {
            const int m = nt - k - 1;
            sv.launch("step", k_step_tiles<F>, m * (m + 1) / 2, 256, A, ld, k, Dinv, d.st);
        }
        if (nt > 1) sv.launch("panel_down", k_panel_down<F>, nt * (nt - 1) / 2, 256, A, ld, (const LMState*)d.st);
        sv.launch("block_inverse", k_block_inverse<F>, nb * kSolveBlk, 256, (const F*)A, ld, (const F*)Dinv, Binv, BinvT,
                  nt, (const LMState*)d.st);
        solve<F>(sv, A, Binv, BinvT, (const double*)d.bc, d.px, 0);
        for (int it = 0; it < refine; ++it) {
            sv.launch("residual", k_residual, div_up(n, kWaves), kWG, (const double*)S.p, ld, (const double*)d.bc,
                      (const double*)d.px, r.p, n, (const LMState*)d.st);
            solve<F>(sv, A, Binv, BinvT, (const double*)r.p, d.px, 1);
        }
    }

    template <typename SolverT>
    void enqueue_solve(SolverT& sv) {
        auto& d = sv.d;
        const int nc = d.nc;
        sv.launch("dense_store",
                  pick_model(d.model, k_dense_store<Real, kModelSnavely>, k_dense_store<Real, kModelSnavelyG2o>,
                             k_dense_store<Real, kModelStereo>),
                  xcd_grid(nc), kWG, d, js.p);
        const size_t nn = (size_t)ld * ld;
        sv.launch("fill_S", k_fill_S, (int)((nn + 255) / 256), 256, S.p, n, ld, (const LMState*)d.st);
        if (nblocks)
            sv.launch("schur_pairs", k_schur_blocks<Real>, nblocks, 64, d, (const Real*)js.p, (const int*)blk_a.p,
                      (const int*)blk_b.p, (const int*)blk_off.p, (const int2*)pairs.p, S.p, ld, nblocks);
        if (f64) {
            factor_and_solve<double>(sv, S.p, Dinvd.p, Binvd.p, BinvTd.p);
        } else {
            sv.launch("convert_S", k_convert_S<float>, (int)((nn + 255) / 256), 256, (const double*)S.p, Af.p, nn,
                      (const LMState*)d.st);
            factor_and_solve<float>(sv, Af.p, Dinvf.p, Binvf.p, BinvTf.p);
        }
        sv.launch("set_iters", k_set_linear_iters, 1, 1, d.st, 1 + refine);
    }
};

}  // namespace mba
