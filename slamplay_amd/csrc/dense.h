// dense.h -- explicit reduced camera system and its dense Cholesky solve (the DENSE_SCHUR /
// SPARSE_SCHUR step of test_bundle_adjustment_ceres.cpp:97 for problems up to a few hundred cameras).
//
//   S = U* - sum_points sum_{i,j in track} W~_i Vinv_p W~_j'     (9 nc x 9 nc, FP64)
// with U*'s per-observation part folded into the (i, i) pairs: S_aa = D_a^2 + sum J~'(I - G)J~ - ...
// is assembled without atomics: the host enumerates every ordered observation pair (i, j) of a
// point with camera(i) <= camera(j) once per problem and groups the pairs by camera block (a, b);
// one workgroup owns one 9x9 block and reduces its pair list in a fixed order.  A pair contributes
// Jc~_i' (Hp_i Jp~_j') Jc~_j with Hp_i = Jp~_i Vinv_p, i.e. a 2x2 middle factor.
//
// The factorization is a tiled right-looking Cholesky (32 x 32 tiles): per step, the diagonal
// tile is factored in LDS and its triangular inverse kept, the panel below is multiplied by that
// inverse (TRSM as a small GEMM) and the trailing lower tiles are updated (SYRK/GEMM).  The
// triangular solves use the kept tile inverses, so no step has a serial dependency chain longer
// than one tile.  In mixed precision the factor is FP32 and the step is refined in FP64 against
// the FP64 S (refine_steps iterations of r = b - S y, y += S_f^-1 r).
#pragma once

#include <algorithm>
#include <cmath>

#include "lm_kernels.h"
#include "solver.h"

namespace mba {

constexpr int kTile = 32;
constexpr int kJs = 36;  // per-observation stride of Jc~ (18), Jp~ (6), Hp (6), I - G (3)

template <typename Real>
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
        corrected_jacobian<Real>(d, P, j, p, r, Jc, Jp);
        Real* o = js + kJs * (size_t)j;
        for (int i = 0; i < 9; ++i) {
            o[i] = Jc[i] * sc[i];
            o[9 + i] = Jc[9 + i] * sc[i];
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

template <typename Real>
__global__ void __launch_bounds__(kWG) k_schur_pairs(Dev<Real> d, const Real* __restrict__ js,
                                                     const int* __restrict__ blk_a, const int* __restrict__ blk_b,
                                                     const int* __restrict__ blk_off, const int* __restrict__ pair_i,
                                                     const int* __restrict__ pair_j, double* __restrict__ S, int ld) {
    LMState* st = d.st;
    if (st->done || st->step_invalid) return;
    __shared__ Real A[32][18];
    __shared__ Real T[32][18];
    const int b = blockIdx.x, t = threadIdx.x;
    const int ca = blk_a[b], cb = blk_b[b];
    const int r = t / 9, cc = t % 9;
    double acc = 0.0;
    for (int base = blk_off[b]; base < blk_off[b + 1]; base += 32) {
        const int n = min(32, blk_off[b + 1] - base);
        if (t < n) {
            const Real* Ji = js + kJs * (size_t)pair_i[base + t];
            const Real* Jj = js + kJs * (size_t)pair_j[base + t];
            const Real* Hp = Ji + 24;
            const Real* Jp = Jj + 18;
            Real G00 = Hp[0] * Jp[0] + Hp[1] * Jp[1] + Hp[2] * Jp[2];
            Real G01 = Hp[0] * Jp[3] + Hp[1] * Jp[4] + Hp[2] * Jp[5];
            Real G10 = Hp[3] * Jp[0] + Hp[4] * Jp[1] + Hp[5] * Jp[2];
            Real G11 = Hp[3] * Jp[3] + Hp[4] * Jp[4] + Hp[5] * Jp[5];
            if (pair_i[base + t] == pair_j[base + t]) {
                // an observation with itself: its J~c'J~c share of U* folded in, J~c'(I - G)J~c
                // with I - G formed in FP64 (see point_complement); negated as acc is subtracted
                G00 = -Ji[30];
                G01 = G10 = -Ji[31];
                G11 = -Ji[32];
            }
            for (int k = 0; k < 9; ++k) {
                T[t][k] = G00 * Jj[k] + G01 * Jj[9 + k];
                T[t][9 + k] = G10 * Jj[k] + G11 * Jj[9 + k];
            }
            for (int k = 0; k < 18; ++k) A[t][k] = Ji[k];
        }
        __syncthreads();
        if (t < 81)
            for (int k = 0; k < n; ++k) acc += (double)(A[k][r] * T[k][cc] + A[k][9 + r] * T[k][9 + cc]);
        __syncthreads();
    }
    if (t < 81) {
        const double v = (ca == cb && r == cc ? d.dc2[9 * (size_t)ca + r] : 0.0) - acc;
        S[(size_t)(9 * ca + r) * ld + 9 * cb + cc] = v;
        if (ca != cb) S[(size_t)(9 * cb + cc) * ld + 9 * ca + r] = v;
    }
}

template <typename F>
__global__ void k_convert_S(const double* __restrict__ S, F* __restrict__ A, size_t n, const LMState* st) {
    if (st->done || st->step_invalid) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[i] = (F)S[i];
}

// Factor diagonal tile k in LDS; keep its lower-triangular inverse in Dinv[k].  Column j: every
// thread reads the pivot (after the previous barrier), scales its row of column j, and after one
// barrier the trailing triangle is updated; the inverse is formed row by row in LDS (row r of
// L^-1 from rows < r), one barrier per row.
template <typename F>
__global__ void __launch_bounds__(kWG) k_potrf_tile(F* __restrict__ A, int ld, int k, F* __restrict__ Dinv, LMState* st) {
    if (st->done || st->step_invalid) return;
    __shared__ F T[kTile][kTile + 1];
    __shared__ F X[kTile][kTile + 1];
    __shared__ int bad;
    const int t = threadIdx.x;
    const size_t o = (size_t)(kTile * k) * ld + kTile * k;
    for (int idx = t; idx < kTile * kTile; idx += kWG) T[idx / kTile][idx % kTile] = A[o + (size_t)(idx / kTile) * ld + idx % kTile];
    if (t == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < kTile; ++j) {
        const F dj = T[j][j];
        const bool nonpos = !(dj > (F)0) || !isfinite((double)dj);
        const F ljj = sqrt(nonpos ? (F)1 : dj), inv = (F)1 / ljj;
        if (t > j && t < kTile) T[t][j] *= inv;
        __syncthreads();
        // T[j][j] is not read again in this step: the update reads column j below the diagonal only
        if (t == 0) {
            T[j][j] = ljj;
            if (nonpos) bad = 1;
        }
        for (int idx = t; idx < kTile * kTile; idx += kWG) {
            const int r = idx / kTile, c = idx % kTile;
            if (c > j && c <= r) T[r][c] -= T[r][j] * T[c][j];
        }
        __syncthreads();
    }
    if (bad) {
        if (t == 0) st->step_invalid = 1;
        return;
    }
    for (int idx = t; idx < kTile * kTile; idx += kWG) {
        const int r = idx / kTile, c = idx % kTile;
        if (c <= r) A[o + (size_t)r * ld + c] = T[r][c];
    }
    // X = L^-1, row r from the rows above it (thread c owns column c)
    for (int r = 0; r < kTile; ++r) {
        if (t < kTile) {
            const int c = t;
            F v = (r == c) ? (F)1 : (F)0;
            for (int m = c; m < r; ++m) v -= T[r][m] * X[m][c];
            X[r][c] = r < c ? (F)0 : v / T[r][r];
        }
        __syncthreads();
    }
    for (int idx = t; idx < kTile * kTile; idx += kWG) Dinv[(size_t)k * kTile * kTile + idx] = X[idx / kTile][idx % kTile];
}

// Panel: A_ik <- A_ik L_kk^-T for i > k (one WG per tile row).
template <typename F>
__global__ void __launch_bounds__(kWG) k_trsm_tiles(F* __restrict__ A, int ld, int k, const F* __restrict__ Dinv, const LMState* st) {
    if (st->done || st->step_invalid) return;
    __shared__ F Li[kTile][kTile + 1];
    __shared__ F X[kTile][kTile + 1];
    const int i = k + 1 + blockIdx.x, t = threadIdx.x;
    const size_t o = (size_t)(kTile * i) * ld + kTile * k;
    for (int idx = t; idx < kTile * kTile; idx += kWG) {
        Li[idx / kTile][idx % kTile] = Dinv[(size_t)k * kTile * kTile + idx];
        X[idx / kTile][idx % kTile] = A[o + (size_t)(idx / kTile) * ld + idx % kTile];
    }
    __syncthreads();
    for (int idx = t; idx < kTile * kTile; idx += kWG) {
        const int r = idx / kTile, c = idx % kTile;
        F v = 0;
        for (int m = 0; m <= c; ++m) v += X[r][m] * Li[c][m];
        A[o + (size_t)r * ld + c] = v;
    }
}

// Trailing update: A_ij -= A_ik A_jk' for k < j <= i (one WG per lower tile).
template <typename F>
__global__ void __launch_bounds__(kWG) k_syrk_tiles(F* __restrict__ A, int ld, int k, const LMState* st) {
    if (st->done || st->step_invalid) return;
    __shared__ F Ai[kTile][kTile + 1];
    __shared__ F Aj[kTile][kTile + 1];
    const int b = blockIdx.x, t = threadIdx.x;
    int ii = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((ii + 1) * (ii + 2) / 2 <= b) ++ii;
    while (ii * (ii + 1) / 2 > b) --ii;
    const int jj = b - ii * (ii + 1) / 2;
    const int i = k + 1 + ii, j = k + 1 + jj;
    const size_t oi = (size_t)(kTile * i) * ld + kTile * k, oj = (size_t)(kTile * j) * ld + kTile * k;
    for (int idx = t; idx < kTile * kTile; idx += kWG) {
        Ai[idx / kTile][idx % kTile] = A[oi + (size_t)(idx / kTile) * ld + idx % kTile];
        Aj[idx / kTile][idx % kTile] = A[oj + (size_t)(idx / kTile) * ld + idx % kTile];
    }
    __syncthreads();
    const size_t od = (size_t)(kTile * i) * ld + kTile * j;
    for (int idx = t; idx < kTile * kTile; idx += kWG) {
        const int r = idx / kTile, c = idx % kTile;
        F v = 0;
        for (int m = 0; m < kTile; ++m) v += Ai[r][m] * Aj[c][m];
        A[od + (size_t)r * ld + c] -= v;
    }
}

// Solve (L L') x = b with the tiled factor (single WG); x = y + solution when accumulate != 0.
template <typename F>
__global__ void __launch_bounds__(1024) k_chol_solve(const F* __restrict__ A, int ld, const F* __restrict__ Dinv,
                                                     const double* __restrict__ b, double* __restrict__ x, int n,
                                                     int accumulate, LMState* st) {
    if (st->done || st->step_invalid) return;
    extern __shared__ double z[];
    __shared__ double zt[kTile];
    const int t = threadIdx.x, nt = ld / kTile;
    for (int i = t; i < ld; i += blockDim.x) z[i] = i < n ? b[i] : 0.0;
    __syncthreads();
    for (int k = 0; k < nt; ++k) {  // forward: L z = b
        if (t < kTile) {
            double v = 0.0;
            for (int m = 0; m <= t; ++m) v += (double)Dinv[(size_t)k * kTile * kTile + t * kTile + m] * z[kTile * k + m];
            zt[t] = v;
        }
        __syncthreads();
        if (t < kTile) z[kTile * k + t] = zt[t];
        __syncthreads();
        for (int i = kTile * (k + 1) + t; i < ld; i += blockDim.x) {
            double v = 0.0;
            for (int m = 0; m < kTile; ++m) v += (double)A[(size_t)i * ld + kTile * k + m] * z[kTile * k + m];
            z[i] -= v;
        }
        __syncthreads();
    }
    for (int k = nt - 1; k >= 0; --k) {  // backward: L' x = z
        if (t < kTile) {
            double v = 0.0;
            for (int m = t; m < kTile; ++m) v += (double)Dinv[(size_t)k * kTile * kTile + m * kTile + t] * z[kTile * k + m];
            zt[t] = v;
        }
        __syncthreads();
        if (t < kTile) z[kTile * k + t] = zt[t];
        __syncthreads();
        for (int i = t; i < kTile * k; i += blockDim.x) {
            double v = 0.0;
            for (int m = 0; m < kTile; ++m) v += (double)A[(size_t)(kTile * k + m) * ld + i] * z[kTile * k + m];
            z[i] -= v;
        }
        __syncthreads();
    }
    for (int i = t; i < n; i += blockDim.x) x[i] = accumulate ? x[i] + z[i] : z[i];
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

template <typename Real>
struct DenseSchur {
    int n = 0, ld = 0, nt = 0, nblocks = 0, refine = 0;
    DevBuf<Real> js;
    DevBuf<int> blk_a, blk_b, blk_off, pair_i, pair_j;
    DevBuf<double> S, r;
    DevBuf<float> Af, Dinvf;
    DevBuf<double> Dinvd;
    int64_t device_bytes() const {
        return js.bytes() + blk_a.bytes() * 3 + pair_i.bytes() * 2 + S.bytes() + r.bytes() + Af.bytes() + Dinvf.bytes() +
               Dinvd.bytes();
    }

    int create(const ba_problem_view* v, const Layout& L, const ba_options& o, hipStream_t s) {
        (void)v;
        const int nc = L.nc;
        n = 9 * nc;
        nt = std::max(1, (n + kTile - 1) / kTile);
        ld = nt * kTile;
        refine = sizeof(Real) == sizeof(double) ? 0 : std::max(0, o.refine_steps);
        PairBlocks pb;
        build_pair_blocks(L, &pb);
        const std::vector<int>&ba = pb.a, &bb = pb.b, &boff = pb.off, &pi = pb.i, &pj = pb.j;
        nblocks = (int)ba.size();
        BA_TRY(blk_a.upload(ba.data(), ba.size(), s));
        BA_TRY(blk_b.upload(bb.data(), bb.size(), s));
        BA_TRY(blk_off.upload(boff.data(), boff.size(), s));
        BA_TRY(pair_i.upload(pi.data(), pi.size(), s));
        BA_TRY(pair_j.upload(pj.data(), pj.size(), s));
        BA_TRY(js.alloc(kJs * (size_t)std::max(L.no, 1)));
        BA_TRY(S.alloc((size_t)ld * ld));
        BA_TRY(r.alloc(ld));
        if (sizeof(Real) == sizeof(double)) {
            BA_TRY(Dinvd.alloc((size_t)nt * kTile * kTile));
        } else {
            BA_TRY(Af.alloc((size_t)ld * ld));
            BA_TRY(Dinvf.alloc((size_t)nt * kTile * kTile));
        }
        BA_HIP(hipStreamSynchronize(s));
        return BA_OK;
    }

    template <typename F, typename SolverT>
    void factor_and_solve(SolverT& sv, F* A, F* Dinv) {
        auto& d = sv.d;
        for (int k = 0; k < nt; ++k) {
            sv.launch("potrf", k_potrf_tile<F>, 1, kWG, A, ld, k, Dinv, d.st);
            if (k + 1 < nt) {
                const int m = nt - k - 1;
                sv.launch("trsm", k_trsm_tiles<F>, m, kWG, A, ld, k, (const F*)Dinv, (const LMState*)d.st);
                sv.launch("syrk", k_syrk_tiles<F>, m * (m + 1) / 2, kWG, A, ld, k, (const LMState*)d.st);
            }
        }
        const size_t shm = (size_t)ld * sizeof(double);
        sv.launch_shm("chol_solve", k_chol_solve<F>, 1, 1024, shm, (const F*)A, ld, (const F*)Dinv, (const double*)d.bc,
                      d.px, n, 0, d.st);
        for (int it = 0; it < refine; ++it) {
            sv.launch("residual", k_residual, div_up(n, kWaves), kWG, (const double*)S.p, ld, (const double*)d.bc,
                      (const double*)d.px, r.p, n, (const LMState*)d.st);
            sv.launch_shm("chol_solve", k_chol_solve<F>, 1, 1024, shm, (const F*)A, ld, (const F*)Dinv,
                          (const double*)r.p, d.px, n, 1, d.st);
        }
    }

    template <typename SolverT>
    void enqueue_solve(SolverT& sv) {
        auto& d = sv.d;
        const int nc = d.nc;
        sv.launch("dense_store", k_dense_store<Real>, xcd_grid(nc), kWG, d, js.p);
        const size_t nn = (size_t)ld * ld;
        sv.launch("fill_S", k_fill_S, (int)((nn + 255) / 256), 256, S.p, n, ld, (const LMState*)d.st);
        sv.launch("schur_pairs", k_schur_pairs<Real>, nblocks, kWG, d, (const Real*)js.p, (const int*)blk_a.p,
                  (const int*)blk_b.p, (const int*)blk_off.p, (const int*)pair_i.p, (const int*)pair_j.p, S.p, ld);
        if (sizeof(Real) == sizeof(double)) {
            factor_and_solve<double>(sv, S.p, Dinvd.p);
        } else {
            sv.launch("convert_S", k_convert_S<float>, (int)((nn + 255) / 256), 256, (const double*)S.p, Af.p, nn,
                      (const LMState*)d.st);
            factor_and_solve<float>(sv, Af.p, Dinvf.p);
        }
        sv.launch("set_iters", k_set_linear_iters, 1, 1, d.st, 1 + refine);
    }
};

}  // namespace mba
