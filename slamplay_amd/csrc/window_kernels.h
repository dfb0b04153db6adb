// window_kernels.h -- batched local bundle adjustment: one workgroup runs whole LM iterations of
// one window (the dense-Schur solve of ba_solve, restated for windows of a few dozen cameras).
//
// Per LM iteration, with workgroup barriers between phases:
//   lin     (after an accepted step) FP64 camera prep; waves over cameras evaluate the corrected
//           residuals and Jacobians once, in FP64, into the per-observation j64 records, with
//           diag(J_c'J_c) and J_c'r per camera (wave reductions) and J_p'J_p, J_p'r per
//           observation into slot order; threads over points sum their slots;
//   damp    V*^-1 (FP64), q = V*^-1 g~_p, D_c^2;
//   records per observation the working-precision inputs of the reduced system (J~_c, J~_p,
//           H = J~_p V*^-1, I - G) from j64; reduced gradient b_c = sum J~_c'(r~ - J~_p q) in FP64;
//   schur   each wave takes 9x9 blocks (a, b) of S from a shared counter, largest first, stages
//           the records of 32 observation pairs at a time in LDS with vector loads, and a lane
//           accumulates every 32nd pair of the block for rows 0-4 or 5-8 in registers; the 32
//           partial sums are combined by a butterfly in FP64 and the block is written to S in
//           global memory, which is then copied into LDS;
//   chol    right-looking Cholesky of S in LDS, one column at a time over the whole workgroup;
//   solve   blocked triangular solves (16-row diagonal blocks by one wave, the off-diagonal
//           updates by the workgroup).  Mixed precision: preconditioned CG in FP64 with the FP32
//           factor as preconditioner and the FP64 operator of the LM system applied implicitly
//           through the j64 records (S y = D_c^2 y + sum J~_c'(J~_c y - J~_p z),
//           z = V*^-1 sum J~_p'J~_c y), until the relative residual is 1e-12: the step is the FP64
//           step also at trust radii where cond(S) is beyond what an FP32 factor resolves;
//   back    u = J~_p'(r~ - J~_c y_c) per observation (FP64), y_p = V*^-1 sum u per point,
//           candidate parameters;
//   decide  model cost change from j64, FP64 candidate cost, Ceres LM rule.
// A window's arithmetic and reduction order depend only on the window -- per-camera sums are
// combined in camera order, per-point sums by one fixed wave, S blocks by one wave each in a fixed
// lane order -- so its result is bitwise the same whichever other windows share the launch and at
// either workgroup size.
#pragma once

#include "window.h"

namespace mba {

// Orders the LDS accesses of one wave around an exchange through its private staging area (the
// wave's LDS operations complete in order; the fences keep the compiler from moving them).
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NT>
__device__ __forceinline__ double wg_max(double v, double* red) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_max(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double m = red[0];
    for (int w = 1; w < NW; ++w) m = fmax(m, red[w]);
    __syncthreads();
    return m;
}

constexpr int kWinMaxN = 288;  // 9 * 32 cameras
constexpr int kWinMaxK = kWinMaxN / 9;
constexpr int kWinTb = 16;     // diagonal block of the triangular solves

struct WinPtrs {
    int No, n_slots;
    const int *cam_off, *co_pt, *slot_of, *chunk_len, *chunk_off;
    const int4* blk;
    const int2* pairs;
    const double* co_uv64;
    double *cam_x, *cam_c, *udiag, *gc, *sc, *dc2, *bc, *V, *gp, *sp, *Vinv, *yp, *j64;
    double4 *pt_x, *pt_c, *pv, *qp;
    CameraPrep<double>*prep, *prep_c;
    WinState* st;
};

__device__ __forceinline__ WinPtrs win_ptrs(const WinDesc& D, char* arena) {
    WinPtrs w;
    w.No = D.No;
    w.n_slots = D.n_slots;
    w.cam_off = at<int>(arena, D.o_cam_off);
    w.co_pt = at<int>(arena, D.o_co_pt);
    w.slot_of = at<int>(arena, D.o_slot_of);
    w.chunk_len = at<int>(arena, D.o_chunk_len);
    w.chunk_off = at<int>(arena, D.o_chunk_off);
    w.blk = at<int4>(arena, D.o_blk);
    w.pairs = at<int2>(arena, D.o_pairs);
    w.co_uv64 = at<double>(arena, D.o_co_uv64);
    w.cam_x = at<double>(arena, D.o_cam_x);
    w.cam_c = at<double>(arena, D.o_cam_c);
    w.udiag = at<double>(arena, D.o_udiag);
    w.gc = at<double>(arena, D.o_gc);
    w.sc = at<double>(arena, D.o_sc);
    w.dc2 = at<double>(arena, D.o_dc2);
    w.bc = at<double>(arena, D.o_bc);
    w.V = at<double>(arena, D.o_V);
    w.gp = at<double>(arena, D.o_gp);
    w.sp = at<double>(arena, D.o_sp);
    w.Vinv = at<double>(arena, D.o_Vinv);
    w.yp = at<double>(arena, D.o_yp);
    w.j64 = at<double>(arena, D.o_j64);
    w.pt_x = at<double4>(arena, D.o_pt_x);
    w.pt_c = at<double4>(arena, D.o_pt_c);
    w.pv = at<double4>(arena, D.o_pv);
    w.qp = at<double4>(arena, D.o_qp);
    w.prep = at<CameraPrep<double>>(arena, D.o_prep);
    w.prep_c = at<CameraPrep<double>>(arena, D.o_prep_c);
    w.st = at<WinState>(arena, D.o_state);
    return w;
}

// Component e of the FP64 record of observation j (stored component-major: j64[e][j], so the
// lanes of a wave, which own consecutive observations, read consecutive addresses).
__device__ __forceinline__ double j64_at(const WinPtrs& w, int e, int j) { return w.j64[(size_t)e * w.No + j]; }

// Scaled FP64 Jacobians of observation j: J~_c = J_c diag(s_c), J~_p = J_p diag(s_p).
__device__ __forceinline__ void win_scaled(const WinPtrs& w, int j, const double sc[9], const double* sp, double Jc[18],
                                           double Jp[6]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Jc[i] = j64_at(w, i, j) * sc[i];
        Jc[9 + i] = j64_at(w, 9 + i, j) * sc[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Jp[i] = j64_at(w, 18 + i, j) * sp[i];
        Jp[3 + i] = j64_at(w, 21 + i, j) * sp[i];
    }
}

// Slot records: pv holds three double4 arrays of n_slots entries each (component-major).
__device__ __forceinline__ double4& pv_at(const WinPtrs& w, int k, int slot) {
    return w.pv[(size_t)k * w.n_slots + slot];
}

// sum over the SELL slots of point p of the first slot record (x, y, z), in slot order.
__device__ __forceinline__ void win_slot_sum3(const WinPtrs& w, int p, double t[3]) {
    const int ch = p >> 6, l = p & 63, len = w.chunk_len[ch];
    const int base = w.chunk_off[ch];
    t[0] = t[1] = t[2] = 0.0;
    for (int r = 0; r < len; ++r) {
        const double4 u = pv_at(w, 0, base + 64 * r + l);
        t[0] += u.x;
        t[1] += u.y;
        t[2] += u.z;
    }
}

constexpr int kWinPairsPerStage = 32;

// LDS bytes of the per-wave staging area of win_schur: the records of 32 pairs (i and j).
template <typename Real>
constexpr int win_stage_bytes() {
    return 2 * kWinPairsPerStage * kWinJs * (int)sizeof(Real);
}

// 9x9 blocks of the reduced camera system into S (n x n row-major, both triangles, global memory):
// S_ab = [a == b] D_a^2 - sum_pairs J~_c,i' G_ij J~_c,j with G_ij = H_i J~_p,j' (i != j) and
// G_ii = -(I - G)_i.  Blocks are taken from *counter in list order.  Per stage of 32 pairs the
// wave copies the record chunks the products read into its LDS staging area with 4-element vector
// loads (lanes cover the chunks of consecutive pairs), then lanes 0-31 accumulate rows 0-4 and lanes
// 32-63 rows 5-8 of pair q = lane % 32 in registers; the 32 partial sums of a block are combined by
// a butterfly in FP64.
template <typename Real, typename F>
__device__ __forceinline__ void win_schur(const int4* __restrict__ blk, const int2* __restrict__ pairs,
                                          const Real* __restrict__ js, const double* __restrict__ dc2, F* S, int n,
                                          int n_blocks, int* counter, Real* stage) {
    constexpr int PS = kWinPairsPerStage;
    const int lane = threadIdx.x & 63, q = lane & 31, h = lane >> 5;
    const int rb = 5 * h;
    Real* st_i = stage;
    Real* st_j = stage + PS * kWinJs;
    for (;;) {
        int b = 0;
        if (lane == 0) b = atomicAdd(counter, 1);
        b = __shfl(b, 0, 64);
        if (b >= n_blocks) break;
        const int4 bk = blk[b];
        Real acc[5][9];
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
            for (int c = 0; c < 9; ++c) acc[r][c] = (Real)0;
        for (int base = bk.z; base < bk.w; base += PS) {
            const int m = min(PS, bk.w - base);
            int2 pr = make_int2(0, 0);
            if (lane < m) pr = pairs[base + lane];
            // cooperative copy of the chunks the pair products read: i records chunks 0-4 (J~_c) and
            // 6-8 (H, I - G), j records chunks 0-5 (J~_c, J~_p); slot s of 14 per pair
            constexpr int CI = 8, CJ = 6, CP = CI + CJ;
            const int nslot = m * CP;
            for (int s0 = 0; s0 < nslot; s0 += 64) {
                const int s = s0 + lane;
                const int k = min(s / CP, PS - 1), e = s - CP * (s / CP);
                const int oi = __shfl(pr.x, k, 64), oj = __shfl(pr.y, k, 64);
                if (s < nslot) {
                    const bool is_j = e >= CI;
                    const int chunk = is_j ? e - CI : (e < 5 ? e : e + 1);
                    const int obs = is_j ? oj : oi;
                    const R4<Real> v = reinterpret_cast<const R4<Real>*>(js + kWinJs * (size_t)obs)[chunk];
                    *reinterpret_cast<R4<Real>*>((is_j ? st_j : st_i) + kWinJs * k + 4 * chunk) = v;
                }
            }
            const int pi = __shfl(pr.x, q, 64), pj = __shfl(pr.y, q, 64);
            wave_sync_lds();
            if (q < m) {
                // whole 4-element chunks from LDS (record stride 36 elements keeps 8 consecutive
                // lanes on distinct banks)
                const R4<Real>* Ri = reinterpret_cast<const R4<Real>*>(st_i + kWinJs * q);
                const R4<Real>* Rj = reinterpret_cast<const R4<Real>*>(st_j + kWinJs * q);
                Real ji[20], jj[24], hg[12];
#pragma unroll
                for (int c4 = 0; c4 < 5; ++c4) {
                    const R4<Real> v = Ri[c4];
                    ji[4 * c4] = v.x;
                    ji[4 * c4 + 1] = v.y;
                    ji[4 * c4 + 2] = v.z;
                    ji[4 * c4 + 3] = v.w;
                }
#pragma unroll
                for (int c4 = 0; c4 < 3; ++c4) {
                    const R4<Real> v = Ri[6 + c4];
                    hg[4 * c4] = v.x;
                    hg[4 * c4 + 1] = v.y;
                    hg[4 * c4 + 2] = v.z;
                    hg[4 * c4 + 3] = v.w;
                }
#pragma unroll
                for (int c4 = 0; c4 < 6; ++c4) {
                    const R4<Real> v = Rj[c4];
                    jj[4 * c4] = v.x;
                    jj[4 * c4 + 1] = v.y;
                    jj[4 * c4 + 2] = v.z;
                    jj[4 * c4 + 3] = v.w;
                }
                Real G00, G01, G10, G11;
                if (pi == pj) {
                    // an observation with itself carries its share of U*: J~c'(I - G)J~c, negated
                    // because the accumulated sum is subtracted
                    G00 = -hg[6];
                    G01 = G10 = -hg[7];
                    G11 = -hg[8];
                } else {
                    // H_i = hg[0..5], J~_p,j = jj[18..23]
                    G00 = hg[0] * jj[18] + hg[1] * jj[19] + hg[2] * jj[20];
                    G01 = hg[0] * jj[21] + hg[1] * jj[22] + hg[2] * jj[23];
                    G10 = hg[3] * jj[18] + hg[4] * jj[19] + hg[5] * jj[20];
                    G11 = hg[3] * jj[21] + hg[4] * jj[22] + hg[5] * jj[23];
                }
                Real a0[5], a1[5];
#pragma unroll
                for (int r = 0; r < 5; ++r) {
                    // rows rb + r (the second half's spare fifth row repeats row 8)
                    a0[r] = h ? ji[min(5 + r, 8)] : ji[r];
                    a1[r] = h ? ji[9 + min(5 + r, 8)] : ji[9 + r];
                }
#pragma unroll
                for (int c = 0; c < 9; ++c) {
                    const Real j0 = jj[c], j1 = jj[9 + c];
                    const Real t0 = G00 * j0 + G01 * j1, t1 = G10 * j0 + G11 * j1;
#pragma unroll
                    for (int r = 0; r < 5; ++r) acc[r][c] += a0[r] * t0 + a1[r] * t1;
                }
            }
            wave_sync_lds();
        }
        const int ca = bk.x, cb = bk.y;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                double v = (double)acc[r][c];
#pragma unroll
                for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
                const int row = rb + r;
                if (q == 0 && row < 9) {
                    const double s = (ca == cb && row == c ? dc2[9 * ca + row] : 0.0) - v;
                    S[(size_t)(9 * ca + row) * n + 9 * cb + c] = (F)s;
                    if (ca != cb) S[(size_t)(9 * cb + c) * n + 9 * ca + row] = (F)s;
                }
            }
        }
    }
}

// In-place right-looking Cholesky of the n x n LDS matrix L (lower triangle used), one column per
// step over the whole workgroup; dinv receives 1 / L_jj.  *bad is set when a pivot is not positive
// (the factor is then meaningless and the caller rejects the step).
template <typename F, int NT>
__device__ __forceinline__ void win_cholesky(F* L, int n, F* colv, F* dinv, int* bad) {
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int j = 0; j < n; ++j) {
        F djj = L[(size_t)j * n + j];
        const bool nonpos = !(djj > (F)0);
        if (nonpos) djj = (F)1;
        const F d = sqrt(djj), inv = (F)1 / d;
        for (int i = j + 1 + tid; i < n; i += NT) colv[i] = L[(size_t)i * n + j] * inv;
        __syncthreads();
        if (tid == 0) {
            L[(size_t)j * n + j] = d;
            dinv[j] = inv;
            if (nonpos) *bad = 1;
        }
        for (int i = j + 1 + wave; i < n; i += NW) {
            const F lij = colv[i];
            F* row = L + (size_t)i * n;
            if (lane == 0) row[j] = lij;
            for (int k = j + 1 + lane; k <= i; k += 64) row[k] -= lij * colv[k];
        }
        __syncthreads();
    }
}

// x <- (L L')^-1 x for the LDS vector x (double), L the factor above.  Blocks of kWinTb rows: the
// diagonal block by wave 0 (one row per lane), the rest of the vector by the workgroup.  Expects x
// complete on entry (after a barrier) and leaves it complete.
template <typename F, int NT>
__device__ __forceinline__ void win_chol_solve(const F* L, const F* dinv, int n, double* x) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // forward: L z = x
    for (int kb = 0; kb < n; kb += kWinTb) {
        const int m = min(kWinTb, n - kb);
        if (wave == 0) {
            double v = lane < m ? x[kb + lane] : 0.0;
            for (int c = 0; c < m; ++c) {
                const double zc = __shfl(v, c, 64) * (double)dinv[kb + c];
                if (lane == c) v = zc;
                if (lane > c && lane < m) v -= (double)L[(size_t)(kb + lane) * n + kb + c] * zc;
            }
            if (lane < m) x[kb + lane] = v;
        }
        __syncthreads();
        for (int i = kb + m + tid; i < n; i += NT) {
            const F* row = L + (size_t)i * n + kb;
            double s = 0.0;
            for (int c = 0; c < m; ++c) s += (double)row[c] * x[kb + c];
            x[i] -= s;
        }
        __syncthreads();
    }
    // backward: L' y = z
    const int last = ((n - 1) / kWinTb) * kWinTb;
    for (int kb = last; kb >= 0; kb -= kWinTb) {
        const int m = min(kWinTb, n - kb);
        if (wave == 0) {
            double v = lane < m ? x[kb + lane] : 0.0;
            for (int c = m - 1; c >= 0; --c) {
                const double yc = __shfl(v, c, 64) * (double)dinv[kb + c];
                if (lane == c) v = yc;
                if (lane < c) v -= (double)L[(size_t)(kb + c) * n + kb + lane] * yc;
            }
            if (lane < m) x[kb + lane] = v;
        }
        __syncthreads();
        for (int i = tid; i < kb; i += NT) {
            double s = 0.0;
            for (int c = 0; c < m; ++c) s += (double)L[(size_t)(kb + c) * n + i] * x[kb + c];
            x[i] -= s;
        }
        __syncthreads();
    }
}

// out = S y for the reduced camera system in FP64, applied through the j64 records:
// S y = D_c^2 y + sum_j J~_c,j'(J~_c,j y - J~_p,j z_p(j)), z_p = V*^-1 sum_{j in p} J~_p,j' J~_c,j y.
// y and out are LDS vectors of 9K; the point vectors z go through the slot records and w.yp.
template <int NT>
__device__ __forceinline__ void win_apply_S(const WinPtrs& w, int K, int P, const double* y, double* out) {
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int c = wave; c < K; c += NW) {
        double scc[9], yc[9];
        for (int i = 0; i < 9; ++i) {
            scc[i] = w.sc[9 * (size_t)c + i];
            yc[i] = y[9 * c + i];
        }
        for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
            const int p = w.co_pt[j];
            double Jc[18], Jp[6];
            win_scaled(w, j, scc, w.sp + 3 * (size_t)p, Jc, Jp);
            double t0 = 0.0, t1 = 0.0;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                t0 += Jc[i] * yc[i];
                t1 += Jc[9 + i] * yc[i];
            }
            pv_at(w, 0, w.slot_of[j]) =
                make_double4(Jp[0] * t0 + Jp[3] * t1, Jp[1] * t0 + Jp[4] * t1, Jp[2] * t0 + Jp[5] * t1, 0.0);
        }
    }
    __syncthreads();
    for (int p = tid; p < P; p += NT) {
        double t[3], z[3];
        win_slot_sum3(w, p, t);
        sym3_mul<double>(w.Vinv + 6 * (size_t)p, t, z);
        for (int i = 0; i < 3; ++i) w.yp[3 * (size_t)p + i] = z[i];
    }
    __syncthreads();
    for (int c = wave; c < K; c += NW) {
        double scc[9], yc[9], acc[9];
        for (int i = 0; i < 9; ++i) {
            scc[i] = w.sc[9 * (size_t)c + i];
            yc[i] = y[9 * c + i];
            acc[i] = 0.0;
        }
        for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
            const int p = w.co_pt[j];
            double Jc[18], Jp[6];
            win_scaled(w, j, scc, w.sp + 3 * (size_t)p, Jc, Jp);
            const double* z = w.yp + 3 * (size_t)p;
            double e0 = -(Jp[0] * z[0] + Jp[1] * z[1] + Jp[2] * z[2]);
            double e1 = -(Jp[3] * z[0] + Jp[4] * z[1] + Jp[5] * z[2]);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                e0 += Jc[i] * yc[i];
                e1 += Jc[9 + i] * yc[i];
            }
#pragma unroll
            for (int i = 0; i < 9; ++i) acc[i] += Jc[i] * e0 + Jc[9 + i] * e1;
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] = wave_sum(acc[i]);
        if (lane == 0)
            for (int i = 0; i < 9; ++i) out[9 * c + i] = w.dc2[9 * (size_t)c + i] * yc[i] + acc[i];
    }
    __syncthreads();
}

// a'b over n entries of two LDS vectors, by wave 0 in a fixed order; every thread gets the value.
__device__ __forceinline__ double win_dot(const double* a, const double* b, int n, double* sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave == 0) {
        double s = 0.0;
        for (int k = lane; k < n; k += 64) s += a[k] * b[k];
        s = wave_sum(s);
        if (lane == 0) *sh = s;
    }
    __syncthreads();
    const double v = *sh;
    __syncthreads();
    return v;
}

constexpr double kWinCgTol = 1e-12;  // relative residual at which the mixed-precision CG stops

template <typename Real, int NT>
__global__ void __launch_bounds__(NT) k_window_lm(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                  const int* __restrict__ work, int n_work, int iters, Opts o,
                                                  int refine) {
    using F = Real;  // reduced system and factor: FP32 in mixed precision (refined in FP64), FP64 otherwise
    constexpr int NW = NT / 64;
    extern __shared__ __align__(32) unsigned char smem[];
    __shared__ double red[NW];
    __shared__ double xv[kWinMaxN];        // reduced gradient, then the camera step y_c
    __shared__ double rv[kWinMaxN];        // CG: residual
    __shared__ double zv[kWinMaxN];        // CG: preconditioned residual
    __shared__ double dv[kWinMaxN];        // CG: search direction
    __shared__ double qv[kWinMaxN];        // CG: S times the search direction
    __shared__ double sdot;
    __shared__ F colv[kWinMaxN];           // Cholesky: current column
    __shared__ F dinv[kWinMaxN];           // 1 / diagonal of the factor
    __shared__ double cpart[kWinMaxK][4];  // per-camera partial sums, combined in camera order
    __shared__ int sflag[4];               // 0: invalid point block, 1: accepted / factor failure, 2: block counter
    const int wi = blockIdx.x;
    if (wi >= n_work) return;
    const WinDesc D = descs[work[wi]];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int K = D.K, P = D.P, n = 9 * K;
    const WinPtrs w = win_ptrs(D, arena);
    Real* js = at<Real>(arena, D.o_js);
    F* Sg = at<F>(arena, D.o_S);  // reduced system as formed (unlisted blocks stay zero)
    WinState* st = w.st;
    F* S = reinterpret_cast<F*>(smem);  // its LDS copy, factored in place

    __syncthreads();
    if (st->done) return;
    if (st->first_lin) {
        // iteration 0: the FP64 cost at the initial parameters
        for (int c = tid; c < K; c += NT) {
            CameraPrep<double> Pc;
            camera_prep<double>(w.cam_x + 9 * (size_t)c, Pc);
            w.prep_c[c] = Pc;
        }
        __syncthreads();
        for (int c = wave; c < K; c += NW) {
            double cost = 0.0;
            for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                const double4 x4 = w.pt_x[w.co_pt[j]];
                const double X[3] = {x4.x, x4.y, x4.z};
                double r[2];
                snavely_residual<double>(w.prep_c[c], X, w.co_uv64[2 * (size_t)j], w.co_uv64[2 * (size_t)j + 1], r);
                cost += half_rho(r[0] * r[0] + r[1] * r[1], o.loss, o.loss_a);
            }
            cost = wave_sum(cost);
            if (lane == 0) cpart[c][0] = cost;
        }
        __syncthreads();
        if (tid == 0) {
            double cost = 0.0;
            for (int c = 0; c < K; ++c) cost += cpart[c][0];
            st->x_cost = st->initial_cost = cost;
            st->rows = 1;
            if (!isfinite(cost)) {
                st->done = 1;
                st->termination = BA_FAILURE;
            }
        }
        __syncthreads();
        if (st->done) return;
    }

    for (int it = 0; it < iters; ++it) {
        __syncthreads();
        if (st->done) break;
        // ---------------- lin ----------------
        if (st->need_lin) {
            const int first = st->first_lin;
            for (int c = tid; c < K; c += NT) {
                CameraPrep<double> Pc;
                camera_prep<double>(w.cam_x + 9 * (size_t)c, Pc);
                w.prep[c] = Pc;
            }
            __syncthreads();
            double gm = 0.0;
            for (int c = wave; c < K; c += NW) {
                double acc[18];
#pragma unroll
                for (int i = 0; i < 18; ++i) acc[i] = 0.0;
                for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                    const int p = w.co_pt[j];
                    const double4 x4 = w.pt_x[p];
                    const double X[3] = {x4.x, x4.y, x4.z};
                    double r[2], Jc[18], Jp[6];
                    snavely_residual_jacobian<double>(w.prep[c], X, w.co_uv64[2 * (size_t)j],
                                                      w.co_uv64[2 * (size_t)j + 1], r, Jc, Jp);
                    const double s = huber_sqrt_rho1<double>(r[0] * r[0] + r[1] * r[1], o.loss, o.loss_a);
                    double* rec = w.j64 + j;
                    const size_t ld = (size_t)w.No;
#pragma unroll
                    for (int i = 0; i < 18; ++i) rec[i * ld] = Jc[i] *= s;
#pragma unroll
                    for (int i = 0; i < 6; ++i) rec[(18 + i) * ld] = Jp[i] *= s;
                    r[0] *= s;
                    r[1] *= s;
                    rec[24 * ld] = r[0];
                    rec[25 * ld] = r[1];
#pragma unroll
                    for (int i = 0; i < 9; ++i) {
                        acc[i] += Jc[i] * Jc[i] + Jc[9 + i] * Jc[9 + i];
                        acc[9 + i] += Jc[i] * r[0] + Jc[9 + i] * r[1];
                    }
                    const int slot = w.slot_of[j];
                    pv_at(w, 0, slot) = make_double4(Jp[0] * Jp[0] + Jp[3] * Jp[3], Jp[0] * Jp[1] + Jp[3] * Jp[4],
                                                     Jp[0] * Jp[2] + Jp[3] * Jp[5], Jp[1] * Jp[1] + Jp[4] * Jp[4]);
                    pv_at(w, 1, slot) = make_double4(Jp[1] * Jp[2] + Jp[4] * Jp[5], Jp[2] * Jp[2] + Jp[5] * Jp[5],
                                                     Jp[0] * r[0] + Jp[3] * r[1], Jp[1] * r[0] + Jp[4] * r[1]);
                    pv_at(w, 2, slot) = make_double4(Jp[2] * r[0] + Jp[5] * r[1], 0.0, 0.0, 0.0);
                }
#pragma unroll
                for (int i = 0; i < 18; ++i) acc[i] = wave_sum(acc[i]);
                if (lane == 0) {
                    for (int i = 0; i < 9; ++i) {
                        w.udiag[9 * (size_t)c + i] = acc[i];
                        w.gc[9 * (size_t)c + i] = acc[9 + i];
                        gm = fmax(gm, fabs(acc[9 + i]));
                    }
                }
            }
            __syncthreads();
            for (int p = tid; p < P; p += NT) {
                const int ch = p >> 6, l = p & 63, len = w.chunk_len[ch];
                const size_t base = (size_t)w.chunk_off[ch];
                double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int r = 0; r < len; ++r) {
                    const int slot = (int)base + 64 * r + l;
                    const double4 x = pv_at(w, 0, slot), y = pv_at(w, 1, slot), z = pv_at(w, 2, slot);
                    a[0] += x.x; a[1] += x.y; a[2] += x.z; a[3] += x.w;
                    a[4] += y.x; a[5] += y.y; a[6] += y.z; a[7] += y.w;
                    a[8] += z.x;
                }
                for (int i = 0; i < 6; ++i) w.V[6 * (size_t)p + i] = a[i];
                for (int i = 0; i < 3; ++i) w.gp[3 * (size_t)p + i] = a[6 + i];
                gm = fmax(gm, fmax(fabs(a[6]), fmax(fabs(a[7]), fabs(a[8]))));
                if (first) {
                    const double dg[3] = {a[0], a[3], a[5]};
                    for (int i = 0; i < 3; ++i) w.sp[3 * (size_t)p + i] = o.jacobi_scaling ? 1.0 / (1.0 + sqrt(dg[i])) : 1.0;
                }
            }
            if (first)
                for (int k = tid; k < n; k += NT) w.sc[k] = o.jacobi_scaling ? 1.0 / (1.0 + sqrt(w.udiag[k])) : 1.0;
            gm = wg_max<NT>(gm, red);
            if (tid == 0) {
                st->gmax = gm;
                st->need_lin = 0;
                st->first_lin = 0;
                if (first && st->iteration >= o.max_iterations) {
                    st->done = 1;
                    st->termination = BA_NO_CONVERGENCE;
                }
                if (!st->done && gm <= o.gtol) {
                    st->done = 1;
                    st->termination = BA_CONVERGENCE;
                }
                if (!st->done && st->radius <= o.min_radius) {
                    st->done = 1;
                    st->termination = BA_CONVERGENCE;
                }
            }
            __syncthreads();
            if (st->done) break;
        }
        // ---------------- damp ----------------
        if (tid == 0) {
            sflag[0] = 0;
            sflag[1] = 0;
            sflag[2] = 0;
        }
        __syncthreads();
        const double radius = st->radius;
        for (int p = tid; p < P; p += NT) {
            const double* V = w.V + 6 * (size_t)p;
            const double s[3] = {w.sp[3 * (size_t)p], w.sp[3 * (size_t)p + 1], w.sp[3 * (size_t)p + 2]};
            double A[9];
            A[0] = s[0] * V[0] * s[0];
            A[1] = A[3] = s[0] * V[1] * s[1];
            A[2] = A[6] = s[0] * V[2] * s[2];
            A[4] = s[1] * V[3] * s[1];
            A[5] = A[7] = s[1] * V[4] * s[2];
            A[8] = s[2] * V[5] * s[2];
            for (int i = 0; i < 3; ++i) A[4 * i] += fmin(fmax(A[4 * i], o.min_diag), o.max_diag) / radius;
            if (!chol_inplace<3>(A)) {
                sflag[0] = 1;
                continue;
            }
            double Ai[9];
            chol_inverse<3>(A, Ai);
            const double sym[6] = {Ai[0], Ai[1], Ai[2], Ai[4], Ai[5], Ai[8]};
            for (int i = 0; i < 6; ++i) w.Vinv[6 * (size_t)p + i] = sym[i];
            const double g[3] = {s[0] * w.gp[3 * (size_t)p], s[1] * w.gp[3 * (size_t)p + 1], s[2] * w.gp[3 * (size_t)p + 2]};
            w.qp[p] = make_double4(Ai[0] * g[0] + Ai[1] * g[1] + Ai[2] * g[2], Ai[3] * g[0] + Ai[4] * g[1] + Ai[5] * g[2],
                                   Ai[6] * g[0] + Ai[7] * g[1] + Ai[8] * g[2], 0.0);
        }
        for (int k = tid; k < n; k += NT)
            w.dc2[k] = fmin(fmax(w.sc[k] * w.udiag[k] * w.sc[k], o.min_diag), o.max_diag) / radius;
        __syncthreads();
        const bool invalid_blocks = sflag[0] != 0;
        // ---------------- records + reduced gradient ----------------
        if (!invalid_blocks) {
            for (int c = wave; c < K; c += NW) {
                double scc[9];
                for (int i = 0; i < 9; ++i) scc[i] = w.sc[9 * (size_t)c + i];
                double bacc[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) bacc[i] = 0.0;
                for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                    const int p = w.co_pt[j];
                    double Jc[18], Jp[6];
                    win_scaled(w, j, scc, w.sp + 3 * (size_t)p, Jc, Jp);
                    const double* Vi = w.Vinv + 6 * (size_t)p;
                    const double j0[3] = {Jp[0], Jp[1], Jp[2]}, j1[3] = {Jp[3], Jp[4], Jp[5]};
                    double h0[3], h1[3];
                    sym3_mul<double>(Vi, j0, h0);
                    sym3_mul<double>(Vi, j1, h1);
                    Real* out = js + kWinJs * (size_t)j;
#pragma unroll
                    for (int i = 0; i < 18; ++i) out[i] = (Real)Jc[i];
#pragma unroll
                    for (int i = 0; i < 6; ++i) out[18 + i] = (Real)Jp[i];
                    for (int i = 0; i < 3; ++i) {
                        out[24 + i] = (Real)h0[i];
                        out[27 + i] = (Real)h1[i];
                    }
                    out[30] = (Real)(1.0 - (j0[0] * h0[0] + j0[1] * h0[1] + j0[2] * h0[2]));
                    out[31] = (Real)(-(j0[0] * h1[0] + j0[1] * h1[1] + j0[2] * h1[2]));
                    out[32] = (Real)(1.0 - (j1[0] * h1[0] + j1[1] * h1[1] + j1[2] * h1[2]));
                    const double4 q = w.qp[p];
                    const double e0 = j64_at(w, 24, j) - (Jp[0] * q.x + Jp[1] * q.y + Jp[2] * q.z);
                    const double e1 = j64_at(w, 25, j) - (Jp[3] * q.x + Jp[4] * q.y + Jp[5] * q.z);
#pragma unroll
                    for (int i = 0; i < 9; ++i) bacc[i] += Jc[i] * e0 + Jc[9 + i] * e1;
                }
#pragma unroll
                for (int i = 0; i < 9; ++i) bacc[i] = wave_sum(bacc[i]);
                if (lane == 0)
                    for (int i = 0; i < 9; ++i) {
                        w.bc[9 * (size_t)c + i] = bacc[i];
                        xv[9 * c + i] = bacc[i];
                    }
            }
            __syncthreads();
            // ---------------- schur + chol ----------------
            win_schur<Real, F>(w.blk, w.pairs, js, w.dc2, Sg, n, D.n_blocks, &sflag[2],
                               reinterpret_cast<Real*>(smem + (size_t)wave * win_stage_bytes<Real>()));
            __syncthreads();
            // factor; in mixed precision a failed FP32 factorization is retried with a diagonal shift
            // (the factor is only the preconditioner of the FP64 CG below)
            for (int attempt = 0;; ++attempt) {
                F shift = (F)0;
                if (attempt > 0) {
                    double dmax = 0.0;
                    for (int k = 0; k < n; ++k) dmax = fmax(dmax, (double)Sg[(size_t)k * n + k]);
                    shift = (F)(dmax * 1e-6 * pow(100.0, attempt - 1));
                }
                for (int idx = tid; idx < n * n; idx += NT) S[idx] = Sg[idx];
                if (tid == 0) sflag[1] = 0;
                __syncthreads();
                if (attempt > 0)
                    for (int k = tid; k < n; k += NT) S[(size_t)k * n + k] += shift;
                __syncthreads();
                win_cholesky<F, NT>(S, n, colv, dinv, &sflag[1]);
                if (sflag[1] == 0 || refine == 0 || attempt == 3) break;
            }
            if (sflag[1] == 0) {
                // ---------------- solve ----------------
                win_chol_solve<F, NT>(S, dinv, n, xv);  // x0 = M^-1 b
                if (refine > 0) {
                    // preconditioned CG on the FP64 operator with M = the factor: converges to the FP64
                    // step also where cond(S) exceeds what an FP32 factor resolves (large trust radii)
                    for (int k = tid; k < n; k += NT) zv[k] = w.bc[k];
                    __syncthreads();
                    const double bb = win_dot(zv, zv, n, &sdot);
                    win_apply_S<NT>(w, K, P, xv, qv);
                    for (int k = tid; k < n; k += NT) rv[k] = zv[k] - qv[k];
                    __syncthreads();
                    double rr = win_dot(rv, rv, n, &sdot);
                    if (rr > kWinCgTol * kWinCgTol * bb) {
                        for (int k = tid; k < n; k += NT) zv[k] = rv[k];
                        __syncthreads();
                        win_chol_solve<F, NT>(S, dinv, n, zv);
                        for (int k = tid; k < n; k += NT) dv[k] = zv[k];
                        __syncthreads();
                        double rz = win_dot(rv, zv, n, &sdot);
                        for (int cg = 0; cg < refine; ++cg) {
                            win_apply_S<NT>(w, K, P, dv, qv);
                            const double dq = win_dot(dv, qv, n, &sdot);
                            if (!(dq > 0.0)) break;
                            const double alpha = rz / dq;
                            for (int k = tid; k < n; k += NT) {
                                xv[k] += alpha * dv[k];
                                rv[k] -= alpha * qv[k];
                            }
                            __syncthreads();
                            rr = win_dot(rv, rv, n, &sdot);
                            if (rr <= kWinCgTol * kWinCgTol * bb) break;
                            for (int k = tid; k < n; k += NT) zv[k] = rv[k];
                            __syncthreads();
                            win_chol_solve<F, NT>(S, dinv, n, zv);
                            const double rz_new = win_dot(rv, zv, n, &sdot);
                            const double beta = rz_new / rz;
                            rz = rz_new;
                            for (int k = tid; k < n; k += NT) dv[k] = zv[k] + beta * dv[k];
                            __syncthreads();
                        }
                    }
                }
            }
        }
        const bool have_step = !invalid_blocks && sflag[1] == 0;
        __syncthreads();
        // ---------------- back-substitution and candidate ----------------
        // |dx|^2, |x|^2, model, cost: per-camera partials (cpart) and one wave over the points, summed
        // in a fixed order so the result does not depend on the workgroup size
        double nrm[2] = {0.0, 0.0};
        if (have_step) {
            for (int c = wave; c < K; c += NW) {
                double scc[9], y[9];
                for (int i = 0; i < 9; ++i) {
                    scc[i] = w.sc[9 * (size_t)c + i];
                    y[i] = xv[9 * c + i];
                }
                for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                    const int p = w.co_pt[j];
                    double Jc[18], Jp[6];
                    win_scaled(w, j, scc, w.sp + 3 * (size_t)p, Jc, Jp);
                    double a0 = j64_at(w, 24, j), a1 = j64_at(w, 25, j);
#pragma unroll
                    for (int i = 0; i < 9; ++i) {
                        a0 -= Jc[i] * y[i];
                        a1 -= Jc[9 + i] * y[i];
                    }
                    pv_at(w, 0, w.slot_of[j]) = make_double4(Jp[0] * a0 + Jp[3] * a1, Jp[1] * a0 + Jp[4] * a1,
                                                             Jp[2] * a0 + Jp[5] * a1, 0.0);
                }
            }
            __syncthreads();
            for (int p = tid; p < P; p += NT) {
                double t[3], y[3];
                win_slot_sum3(w, p, t);
                sym3_mul<double>(w.Vinv + 6 * (size_t)p, t, y);
                const double4 x4 = w.pt_x[p];
                const double xs[3] = {x4.x, x4.y, x4.z};
                double cand[3];
                for (int i = 0; i < 3; ++i) {
                    w.yp[3 * (size_t)p + i] = y[i];
                    cand[i] = xs[i] - w.sp[3 * (size_t)p + i] * y[i];
                }
                w.pt_c[p] = make_double4(cand[0], cand[1], cand[2], 0.0);
            }
            for (int k = tid; k < n; k += NT) w.cam_c[k] = w.cam_x[k] - w.sc[k] * xv[k];
            __syncthreads();
            if (wave == 0) {
                for (int p = lane; p < P; p += 64) {
                    const double4 x4 = w.pt_x[p];
                    const double xs[3] = {x4.x, x4.y, x4.z};
                    for (int i = 0; i < 3; ++i) {
                        const double dl = -w.sp[3 * (size_t)p + i] * w.yp[3 * (size_t)p + i];
                        nrm[0] += dl * dl;
                        nrm[1] += xs[i] * xs[i];
                    }
                }
                nrm[0] = wave_sum(nrm[0]);
                nrm[1] = wave_sum(nrm[1]);
            }
            for (int c = tid; c < K; c += NT) {
                CameraPrep<double> Pc;
                camera_prep<double>(w.cam_c + 9 * (size_t)c, Pc);
                w.prep_c[c] = Pc;
            }
            __syncthreads();
            // model cost change -(J~s)'(r~ + J~s/2) from j64 (step s = -y) and the FP64 candidate cost
            for (int c = wave; c < K; c += NW) {
                double scc[9], y[9];
                for (int i = 0; i < 9; ++i) {
                    scc[i] = w.sc[9 * (size_t)c + i];
                    y[i] = xv[9 * c + i];
                }
                double cm[4] = {0.0, 0.0, 0.0, 0.0};
                if (lane < 9) {
                    const int k = 9 * c + lane;
                    const double dl = -w.sc[k] * xv[k], x = w.cam_x[k];
                    cm[0] = dl * dl;
                    cm[1] = x * x;
                }
                for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                    const int p = w.co_pt[j];
                    double Jc[18], Jp[6];
                    win_scaled(w, j, scc, w.sp + 3 * (size_t)p, Jc, Jp);
                    const double* yp = w.yp + 3 * (size_t)p;
                    double m0 = 0.0, m1 = 0.0;
                    for (int i = 0; i < 9; ++i) {
                        m0 -= Jc[i] * y[i];
                        m1 -= Jc[9 + i] * y[i];
                    }
                    for (int i = 0; i < 3; ++i) {
                        m0 -= Jp[i] * yp[i];
                        m1 -= Jp[3 + i] * yp[i];
                    }
                    cm[2] += m0 * (j64_at(w, 24, j) + 0.5 * m0) + m1 * (j64_at(w, 25, j) + 0.5 * m1);
                    const double4 x4 = w.pt_c[p];
                    const double X[3] = {x4.x, x4.y, x4.z};
                    double r[2];
                    snavely_residual<double>(w.prep_c[c], X, w.co_uv64[2 * (size_t)j], w.co_uv64[2 * (size_t)j + 1], r);
                    cm[3] += half_rho(r[0] * r[0] + r[1] * r[1], o.loss, o.loss_a);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) cm[i] = wave_sum(cm[i]);
                if (lane == 0)
                    for (int i = 0; i < 4; ++i) cpart[c][i] = cm[i];
            }
        }
        __syncthreads();
        // ---------------- decide ----------------
        if (tid == 0) {
            sflag[1] = 0;
            double tot[4] = {nrm[0], nrm[1], 0.0, 0.0};  // wave 0 holds the point sums
            if (have_step)
                for (int c = 0; c < K; ++c)
                    for (int i = 0; i < 4; ++i) tot[i] += cpart[c][i];
            const double model = -tot[2];
            double cand = tot[3];
            const double dx2 = tot[0], x2 = tot[1];
            st->iteration++;
            const bool valid = have_step && model > 0.0 && isfinite(dx2);
            bool record = true;
            if (!valid) {
                if (++st->consecutive_invalid >= o.max_invalid) {
                    st->done = 1;
                    st->termination = BA_FAILURE;
                    record = false;
                } else {
                    st->radius = st->radius / st->decrease_factor;
                    st->decrease_factor *= 2.0;
                    st->num_unsuccessful++;
                }
            } else {
                st->consecutive_invalid = 0;
                if (!isfinite(cand)) cand = DBL_MAX;
                const double step_norm = sqrt(dx2), x_norm = sqrt(x2);
                if (step_norm <= o.ptol * (x_norm + o.ptol)) {
                    st->done = 1;
                    st->termination = BA_CONVERGENCE;
                    record = false;
                } else {
                    const double cost_change = st->x_cost - cand;
                    if (fabs(cost_change) <= o.ftol * st->x_cost) {
                        st->done = 1;
                        st->termination = BA_CONVERGENCE;
                        record = false;
                    } else {
                        const double rho = cost_change / model;
                        if (rho > o.min_rel_decrease) {
                            st->x_cost = cand;
                            const double t = 2.0 * rho - 1.0;
                            st->radius = fmin(o.max_radius, st->radius / fmax(1.0 / 3.0, 1.0 - t * t * t));
                            st->decrease_factor = 2.0;
                            st->need_lin = 1;
                            st->num_successful++;
                            sflag[1] = 1;
                        } else {
                            st->radius = st->radius / st->decrease_factor;
                            st->decrease_factor *= 2.0;
                            st->num_unsuccessful++;
                        }
                    }
                }
            }
            if (record) st->rows++;
            if (record && !st->done) {
                if (st->iteration >= o.max_iterations) {
                    st->done = 1;
                    st->termination = BA_NO_CONVERGENCE;
                } else if (!sflag[1] && st->gmax <= o.gtol) {
                    st->done = 1;
                    st->termination = BA_CONVERGENCE;
                } else if (!sflag[1] && st->radius <= o.min_radius) {
                    st->done = 1;
                    st->termination = BA_CONVERGENCE;
                }
            }
        }
        __syncthreads();
        if (sflag[1]) {
            for (int k = tid; k < n; k += NT) w.cam_x[k] = w.cam_c[k];
            for (int p = tid; p < P; p += NT) w.pt_x[p] = w.pt_c[p];
        }
    }
}

}  // namespace mba
