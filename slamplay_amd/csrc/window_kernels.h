// window_kernels.h -- batched local bundle adjustment: the LM iteration of many windows as a
// fixed sequence of phase kernels (the dense-Schur solve of ba_solve, restated for windows of a
// few dozen cameras).  Each kernel covers every active window at once:
//   k_wlin_cam    (window, camera) FP64 residuals and Jacobians of the camera's observations:
//                 diag(J_c'J_c), J_c'r, J_p'J_p and J_p'r into slot order, and at the start the
//                 FP64 cost (after an accepted step only);
//   k_wlin_pt     (window, point chunk) V and g_p from the slots;
//   k_wdamp       (window) gradient test, Jacobi scales at the start, V*^-1 and q = V*^-1 g~_p
//                 per point, D_c^2;
//   k_wrec_cam    (window, camera) working-precision records of the reduced system (J~_c, J~_p,
//                 H = J~_p V*^-1, I - G) and b_c = sum J~_c'(r~ - J~_p q) in FP64;
//   k_wschur      (window, block group) 9x9 blocks of S, each by one wave from LDS-staged
//                 observation pairs;
//   k_wsolve      (window) S copied as a packed lower triangle into LDS (or the window's global
//                 buffer), right-looking FP64 Cholesky, blocked triangular solves;
//   k_wback_cam   (window, camera) u = J~_p'(r~ - J~_c y_c) into slot order (FP64);
//   k_wback_pt    (window, point chunk) y_p = V*^-1 sum u, candidate points, |dx|^2, |x|^2;
//   k_wcand_cam   (window, camera) candidate camera, model cost change, FP64 candidate cost;
//   k_wdecide     (window) Ceres LM rule (or g2o's Levenberg rule), accepted candidates copied into
//                 the parameters.
// Residual model M (a template parameter of the kernels that evaluate residuals): the Snavely camera
// of the BAL path, or full_slam's stereo rig with the SE3 pose update (stereo.h).
// The per-observation FP64 residual and Jacobian are recomputed by each kernel that needs them
// (k_wlin_cam, k_wrec_cam, k_wback_cam) from the camera's prep, the point and the measurement:
// about 300 FP64 operations against 26 stored doubles written once and read three times per
// iteration; k_wcand_cam evaluates residuals at the candidate only.
// Every kernel reads the window's LM state first and returns when it has nothing to do (or when its
// task lies outside the window it names), so the sequence can be replayed as a graph.  Per-camera and per-chunk sums are written to their own
// slots and combined by the per-window kernels in camera / chunk order, and every 9x9 block of S is
// reduced by one wave in a fixed order: a window's result does not depend on which other windows
// share the launch, nor on the block sizes chosen for the launch.
#pragma once

#include "mfma.h"
#include "pair_block.h"
#include "window.h"

namespace mba {

constexpr int kWinMaxN = 288;  // 9 * 32 cameras
constexpr int kWinMaxK = kWinMaxN / 9;
constexpr int kWinTb = 16;     // diagonal block of the triangular solves
constexpr int kWinCamThreads = 256;  // block size of the per-camera kernels
struct WinPtrs {
    int K, P, No, n_slots;
    const int *cam_off, *co_pt, *slot_of, *chunk_len, *chunk_off;
    const int4* blk;
    const int2* pairs;
    const double* co_uv64;
    double *cam_x, *cam_c, *udiag, *gc, *sc, *dc2, *bc, *yc, *cpart, *ppart;
    double *V, *gp, *sp, *Vinv, *yp;
    double4 *pt_x, *pt_c, *pv, *qp;
    CameraPrep<double>*prep, *prep_c;
    const int* side;       // stereo model: camera side per camera-order observation
    const StereoRig* rig;  // stereo model
    WinState* st;
};

__device__ __forceinline__ WinPtrs win_ptrs(const WinDesc& D, char* arena) {
    WinPtrs w;
    w.K = D.K;
    w.P = D.P;
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
    w.yc = at<double>(arena, D.o_yc);
    w.cpart = at<double>(arena, D.o_cpart);
    w.ppart = at<double>(arena, D.o_ppart);
    w.V = at<double>(arena, D.o_V);
    w.gp = at<double>(arena, D.o_gp);
    w.sp = at<double>(arena, D.o_sp);
    w.Vinv = at<double>(arena, D.o_Vinv);
    w.yp = at<double>(arena, D.o_yp);
    w.pt_x = at<double4>(arena, D.o_pt_x);
    w.pt_c = at<double4>(arena, D.o_pt_c);
    w.pv = at<double4>(arena, D.o_pv);
    w.qp = at<double4>(arena, D.o_qp);
    w.prep = at<CameraPrep<double>>(arena, D.o_prep);
    w.prep_c = at<CameraPrep<double>>(arena, D.o_prep_c);
    w.side = at<int>(arena, D.o_side);
    w.rig = at<StereoRig>(arena, D.o_rig);
    w.st = at<WinState>(arena, D.o_state);
    return w;
}

// The window's rig into the workgroup's copy (stereo model; a no-op for the others).  Thread 0
// copies; the caller's next barrier publishes it.
template <int M>
__device__ __forceinline__ void win_load_rig(const WinPtrs& w, StereoRig& g) {
    if constexpr (M == kModelStereo)
        if (threadIdx.x == 0) g = *w.rig;
}

// Residual and Jacobians of observation j (point p, measurement (u, v)) under model M.
template <int M>
__device__ __forceinline__ void win_model_jacobian(const WinPtrs& w, const CameraPrep<double>& P, const StereoRig& g,
                                                   int j, const double X[3], double r[2], double Jc[18], double Jp[6]) {
    const double u = w.co_uv64[2 * (size_t)j], v = w.co_uv64[2 * (size_t)j + 1];
    if constexpr (M == kModelStereo)
        stereo_residual_jacobian<double>(P, X, g, w.side[j], u, v, r, Jc, Jp);
    else
        snavely_residual_jacobian<double>(P, X, u, v, r, Jc, Jp);
}

template <int M>
__device__ __forceinline__ void win_model_residual(const WinPtrs& w, const CameraPrep<double>& P, const StereoRig& g,
                                                   int j, const double X[3], double r[2]) {
    const double u = w.co_uv64[2 * (size_t)j], v = w.co_uv64[2 * (size_t)j + 1];
    if constexpr (M == kModelStereo)
        stereo_residual<double>(P, X, g, w.side[j], u, v, r);
    else
        snavely_residual<double>(P, X, u, v, r);
}

// Sum of N per-thread values over a workgroup of NT threads in a fixed order (wave butterflies,
// then the waves in order by thread 0); the totals are returned to thread 0 only.
template <int N, int NT>
__device__ __forceinline__ void wg_sum_fixed(double (&v)[N], double* lds) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double s = wave_sum(v[i]);
        if (lane == 0) lds[i * NW + wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < N; ++i) {
            double s = lds[i * NW];
            for (int k = 1; k < NW; ++k) s += lds[i * NW + k];
            v[i] = s;
        }
}

// Huber-corrected FP64 residual and Jacobians of observation j (point p) at the current parameters
// (P: the camera's prep at its current parameters), as k_wlin_cam forms them.
template <int M>
__device__ __forceinline__ void win_obs(const WinPtrs& w, const CameraPrep<double>& P, const StereoRig& g, int j, int p,
                                        const Opts& o, double r[2], double Jc[18], double Jp[6]) {
    const double4 x4 = w.pt_x[p];
    const double X[3] = {x4.x, x4.y, x4.z};
    win_model_jacobian<M>(w, P, g, j, X, r, Jc, Jp);
    const double s = huber_sqrt_rho1<double>(r[0] * r[0] + r[1] * r[1], o.loss, o.loss_a);
#pragma unroll
    for (int i = 0; i < 18; ++i) Jc[i] *= s;
#pragma unroll
    for (int i = 0; i < 6; ++i) Jp[i] *= s;
    r[0] *= s;
    r[1] *= s;
}

// win_obs with the Jacobi scales applied: J~_c = J_c diag(s_c), J~_p = J_p diag(s_p).
template <int M>
__device__ __forceinline__ void win_scaled(const WinPtrs& w, const CameraPrep<double>& P, const StereoRig& g, int j,
                                           int p, const Opts& o, const double sc[9], double r[2], double Jc[18],
                                           double Jp[6]) {
    win_obs<M>(w, P, g, j, p, o, r, Jc, Jp);
    const double* sp = w.sp + 3 * (size_t)p;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Jc[i] *= sc[i];
        Jc[9 + i] *= sc[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Jp[i] *= sp[i];
        Jp[3 + i] *= sp[i];
    }
}

// Slot records: pv holds three double4 arrays of n_slots entries each (component-major).
__device__ __forceinline__ double4& pv_at(const WinPtrs& w, int k, int slot) {
    return w.pv[(size_t)k * w.n_slots + slot];
}

// sum over the SELL slots of point p of the first slot record (x, y, z), in slot order; the loads
// of eight rows are issued together.
__device__ __forceinline__ void win_slot_sum3(const WinPtrs& w, int p, double t[3]) {
    const int ch = p >> 6, l = p & 63, len = w.chunk_len[ch];
    const int base = w.chunk_off[ch];
    t[0] = t[1] = t[2] = 0.0;
    for (int r0 = 0; r0 < len; r0 += 8) {
        double4 u[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) u[i] = pv_at(w, 0, base + 64 * min(r0 + i, len - 1) + l);
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (r0 + i < len) {
                t[0] += u[i].x;
                t[1] += u[i].y;
                t[2] += u[i].z;
            }
    }
}

// The blocks first + stride t, t < count, of one window, taken one at a time through *counter.
template <typename Real>
__device__ __forceinline__ void win_schur(const int4* __restrict__ blk, const int2* __restrict__ pairs,
                                          const Real* __restrict__ js, const double* __restrict__ dc2, double* S, int n,
                                          int first, int stride, int count, int* counter, Real* stage) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __shfl(t, 0, 64);
        if (t >= count) break;
        pair_block<Real, kRecWindow>(blk[first + stride * t], pairs, js, dc2, S, n, stage);
    }
}

// LDS bytes of the per-wave staging area of win_schur.
template <typename Real>
constexpr int win_stage_bytes() {
    return pair_stage_bytes<Real, kRecWindow>();
}

// Packed lower-triangular storage: element (i, j), j <= i, at i (i + 1) / 2 + j (row i contiguous).
__device__ __forceinline__ int tri(int i) { return i * (i + 1) / 2; }

// In-place blocked right-looking Cholesky of the packed n x n matrix L (LDS, or the window's global
// buffer); dinv receives 1 / L_jj.  Panels of kWinNb columns: wave 0 factors the panel's diagonal
// block in registers (lane i holds row i; each column's entries move between lanes through LDS, no
// barrier), then every thread solves its rows of the panel below the block against it
// (x L11' = a, the entries of L11 read from LDS by all threads at once), so a panel costs two
// barriers.  The trailing lower triangle is then updated with the panel by 16 x 16 tiles on the
// matrix cores (one tile per wave at a time, K = 16 in four steps; entries past n or above the
// diagonal are neither read nor written).  *bad is set when a pivot is not positive (the factor is
// then meaningless and the caller rejects the step).  Every entry receives the same sequence of
// operations as in a column-by-column right-looking factorization (its updates by the earlier
// columns in order, then the scaling by 1 / sqrt(pivot)), fixed by the algorithm, not by the
// workgroup size.
constexpr int kWinNb = 16;

template <typename F, int NT>
__device__ __forceinline__ void win_panel(F* L, int n, int jb, int je, F* dinv, F* xch, int* bad) {
    constexpr int NB = kWinNb, RS = (kWinMaxN - NB + NT - 1) / NT;
    const int tid = threadIdx.x, nb = je - jb;
    if (tid < 64) {
        // the diagonal block: lane i < nb holds row jb + i; the pivot's reciprocal square root by
        // the hardware estimate and Newton steps, each column's entries to the other lanes through
        // xch (a wave's LDS operations complete in order)
        const int i = tid;
        F a[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) a[c] = (i < nb && c <= i) ? L[tri(jb + i) + jb + c] : (F)0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < nb) {  // uniform
                F djj = read_lane(a[j], j);
                const bool nonpos = !(djj > (F)0);
                if (nonpos) djj = (F)1;
                const F inv = inv_sqrt(djj), d = djj * inv;
                a[j] = i > j ? a[j] * inv : (i == j ? d : a[j]);
                if (i == 0) {
                    dinv[jb + j] = inv;
                    if (nonpos) *bad = 1;
                }
                if (i < NB) xch[i] = a[j];
                wave_sync_lds();
#pragma unroll
                for (int c = j + 1; c < NB; ++c) {
                    if (c < nb) {  // uniform
                        const F lcj = xch[c];
                        if (i >= c && i < nb) a[c] = fma(-a[j], lcj, a[c]);
                    }
                }
                wave_sync_lds();
            }
        }
        if (i < nb)
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c <= i) L[tri(jb + i) + jb + c] = a[c];
    }
    __syncthreads();
    // the panel's rows below the block: x_j = (a_j - sum_k<j x_k l_jk) / l_jj, k in order
#pragma unroll
    for (int s = 0; s < RS; ++s) {
        const int row = je + tid + NT * s;
        if (row < n) {
            F a[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) a[c] = c < nb ? L[tri(row) + jb + c] : (F)0;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (j < nb) {
#pragma unroll
                    for (int k = 0; k < j; ++k) a[j] = fma(-a[k], L[tri(jb + j) + jb + k], a[j]);
                    a[j] = a[j] * dinv[jb + j];
                }
            }
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < nb) L[tri(row) + jb + c] = a[c];
        }
    }
}

template <typename F, int NT>
__device__ __forceinline__ void win_cholesky(F* L, int n, F* dinv, F* xch, int* bad) {
    constexpr int NW = NT / 64, NB = kWinNb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int jb = 0; jb < n; jb += NB) {
        const int je = min(jb + NB, n);
        win_panel<F, NT>(L, n, jb, je, dinv, xch, bad);
        __syncthreads();
        if (je >= n) break;
        // trailing update: L[r][c] -= sum_{k in [jb, je)} L[r][k] L[c][k] for je <= c <= r < n
        const int m = n - je, T = (m + 15) / 16;
        const int l16 = lane & 15, q = lane >> 4;
        for (int tt = wave; tt < T * (T + 1) / 2; tt += NW) {
            int ti = (int)((sqrtf(8.0f * tt + 1.0f) - 1.0f) * 0.5f);
            while ((ti + 1) * (ti + 2) / 2 <= tt) ++ti;
            while (ti * (ti + 1) / 2 > tt) --ti;
            const int tk = tt - ti * (ti + 1) / 2;
            const int r0 = je + 16 * ti, c0 = je + 16 * tk;
            typename Mfma<F>::acc acc;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = r0 + Mfma<F>::row(lane, e), c = c0 + l16;
                acc[e] = (r < n && c <= r) ? L[tri(r) + c] : (F)0;
            }
            const int ra = r0 + l16, cb = c0 + l16;
#pragma unroll
            for (int s4 = 0; s4 < NB / 4; ++s4) {
                const int k = jb + 4 * s4 + q;
                const F a = (ra < n && k < je) ? L[tri(ra) + k] : (F)0;
                const F b = (cb < n && k < je) ? L[tri(cb) + k] : (F)0;
                acc = Mfma<F>::mma(-a, b, acc);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = r0 + Mfma<F>::row(lane, e), c = c0 + l16;
                if (r < n && c <= r) L[tri(r) + c] = acc[e];
            }
        }
        __syncthreads();
    }
}

// x <- (L L')^-1 x for the LDS vector x (double), L the packed factor above.  Blocks of kWinTb
// rows: the diagonal block by wave 0 (one row per lane), the rest of the vector by the workgroup.
// Expects x complete on entry (after a barrier) and leaves it complete.
template <typename F, int NT>
__device__ __forceinline__ void win_chol_solve(const F* L, const F* dinv, int n, double* x) {
    constexpr int TB = kWinTb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // forward: L z = x
    for (int kb = 0; kb < n; kb += TB) {
        const int m = min(TB, n - kb);
        if (wave == 0) {
            // the lane's row of the diagonal block and the block's 1 / l_cc first, so that the
            // serial chain below waits on no memory
            double lr[TB], dv[TB];
#pragma unroll
            for (int c = 0; c < TB; ++c) {
                lr[c] = (lane < m && c < lane) ? (double)L[tri(kb + lane) + kb + c] : 0.0;
                dv[c] = c < m ? (double)dinv[kb + c] : 0.0;
            }
            double v = lane < m ? x[kb + lane] : 0.0;
#pragma unroll
            for (int c = 0; c < TB; ++c) {
                if (c < m) {
                    const double zc = read_lane(v, c) * dv[c];
                    if (lane == c) v = zc;
                    if (lane > c) v -= lr[c] * zc;
                }
            }
            if (lane < m) x[kb + lane] = v;
        }
        __syncthreads();
        for (int i = kb + m + tid; i < n; i += NT) {
            const F* row = L + tri(i) + kb;
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < TB; ++c)
                if (c < m) s += (double)row[c] * x[kb + c];
            x[i] -= s;
        }
        __syncthreads();
    }
    // backward: L' y = z
    const int last = ((n - 1) / TB) * TB;
    for (int kb = last; kb >= 0; kb -= TB) {
        const int m = min(TB, n - kb);
        if (wave == 0) {
            double lc[TB], dv[TB];
#pragma unroll
            for (int c = 0; c < TB; ++c) {
                lc[c] = (c < m && lane < c) ? (double)L[tri(kb + c) + kb + lane] : 0.0;
                dv[c] = c < m ? (double)dinv[kb + c] : 0.0;
            }
            double v = lane < m ? x[kb + lane] : 0.0;
#pragma unroll
            for (int c = TB - 1; c >= 0; --c) {
                if (c < m) {
                    const double yc = read_lane(v, c) * dv[c];
                    if (lane == c) v = yc;
                    if (lane < c) v -= lc[c] * yc;
                }
            }
            if (lane < m) x[kb + lane] = v;
        }
        __syncthreads();
        for (int i = tid; i < kb; i += NT) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < TB; ++c)
                if (c < m) s += (double)L[tri(kb + c) + i] * x[kb + c];
            x[i] -= s;
        }
        __syncthreads();
    }
}



// ------------------------------------------------------------------------------------------------
// Phase kernels.  ctask: (slot, camera) per workgroup; ptask: (slot, chunk) per 64-thread
// workgroup; wtask: slot per workgroup; stask: (slot, first block, block stride, block count).

template <int NT, int M>
__global__ void __launch_bounds__(NT) k_wlin_cam(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                 const int2* __restrict__ ctask, int ntask, Opts o) {
    __shared__ double lds[19 * (NT / 64)];
    __shared__ CameraPrep<double> Psh;
    __shared__ StereoRig Gsh;
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const int2 tk = ctask[t];
    const WinDesc D = descs[tk.x];
    const WinPtrs w = win_ptrs(D, arena);
    const WinState* st = w.st;
    if (st->done || !st->need_lin || tk.y >= w.K) return;
    const int c = tk.y, first = st->first_lin;
    // the first round's point index and point are requested before the prep and the barrier
    const int j_end = w.cam_off[c + 1];
    int j = w.cam_off[c] + threadIdx.x;
    double4 x4n = make_double4(0.0, 0.0, 0.0, 0.0);
    if (j < j_end) x4n = w.pt_x[w.co_pt[j]];
    if (threadIdx.x == 0) {
        camera_prep<double>(w.cam_x + 9 * (size_t)c, Psh);
        w.prep[c] = Psh;
    }
    win_load_rig<M>(w, Gsh);
    __syncthreads();
    double acc[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) acc[i] = 0.0;  // diag (9), J_c'r (9), cost
    for (; j < j_end; j += NT) {
        const double4 x4 = x4n;
        if (j + NT < j_end) x4n = w.pt_x[w.co_pt[j + NT]];
        const double X[3] = {x4.x, x4.y, x4.z};
        double r[2], Jc[18], Jp[6];
        win_model_jacobian<M>(w, Psh, Gsh, j, X, r, Jc, Jp);
        const double s2 = r[0] * r[0] + r[1] * r[1];
        if (first) acc[18] += half_rho(s2, o.loss, o.loss_a);
        const double s = huber_sqrt_rho1<double>(s2, o.loss, o.loss_a);
#pragma unroll
        for (int i = 0; i < 18; ++i) Jc[i] *= s;
#pragma unroll
        for (int i = 0; i < 6; ++i) Jp[i] *= s;
        r[0] *= s;
        r[1] *= s;
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
    wg_sum_fixed<19, NT>(acc, lds);
    if (threadIdx.x == 0) {
        double gm = 0.0;
        for (int i = 0; i < 9; ++i) {
            w.udiag[9 * (size_t)c + i] = acc[i];
            w.gc[9 * (size_t)c + i] = acc[9 + i];
            gm = fmax(gm, fabs(acc[9 + i]));
        }
        w.cpart[WC_COUNT * (size_t)c + WC_GMAX] = gm;
        w.cpart[WC_COUNT * (size_t)c + WC_COST] = acc[18];
    }
}

// One wave per chunk of 64 points (lane = point).
__global__ void __launch_bounds__(64) k_wlin_pt(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                const int2* __restrict__ ptask, int ntask, Opts o) {
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const int2 tk = ptask[t];
    const WinDesc D = descs[tk.x];
    const WinPtrs w = win_ptrs(D, arena);
    const WinState* st = w.st;
    if (st->done || !st->need_lin || tk.y >= D.n_chunks) return;
    const int ch = tk.y, l = threadIdx.x, p = 64 * ch + l;
    double gm = 0.0;
    if (p < w.P) {
        const int len = w.chunk_len[ch], base = w.chunk_off[ch];
        double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int r0 = 0; r0 < len; r0 += 4) {  // four rows' loads in flight, summed in row order
            double4 x[4], y[4], z[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int slot = base + 64 * min(r0 + i, len - 1) + l;
                x[i] = pv_at(w, 0, slot);
                y[i] = pv_at(w, 1, slot);
                z[i] = pv_at(w, 2, slot);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (r0 + i < len) {
                    a[0] += x[i].x; a[1] += x[i].y; a[2] += x[i].z; a[3] += x[i].w;
                    a[4] += y[i].x; a[5] += y[i].y; a[6] += y[i].z; a[7] += y[i].w;
                    a[8] += z[i].x;
                }
        }
        for (int i = 0; i < 6; ++i) w.V[6 * (size_t)p + i] = a[i];
        for (int i = 0; i < 3; ++i) w.gp[3 * (size_t)p + i] = a[6 + i];
        gm = fmax(fabs(a[6]), fmax(fabs(a[7]), fabs(a[8])));
        if (st->first_lin) {
            const double dg[3] = {a[0], a[3], a[5]};
            for (int i = 0; i < 3; ++i) w.sp[3 * (size_t)p + i] = o.jacobi_scaling ? 1.0 / (1.0 + sqrt(dg[i])) : 1.0;
        }
    }
    gm = wave_max(gm);
    if (l == 0) w.ppart[WP_COUNT * (size_t)ch + WP_GMAX] = gm;
}

// Per window: close the linearization (cost at the start, gradient test, Jacobi scales), then the
// damped point blocks and camera damping of this iteration.
template <int NT>
__global__ void __launch_bounds__(NT) k_wdamp(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                              const int* __restrict__ wtask, int ntask, Opts o) {
    __shared__ int bad;
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const WinDesc D = descs[wtask[t]];
    const WinPtrs w = win_ptrs(D, arena);
    WinState* st = w.st;
    if (st->done) return;
    const int K = w.K, P = w.P, n = 9 * K, tid = threadIdx.x;
    if (st->need_lin) {
        const int first = st->first_lin;
        if (first)
            for (int k = tid; k < n; k += NT) w.sc[k] = o.jacobi_scaling ? 1.0 / (1.0 + sqrt(w.udiag[k])) : 1.0;
        if (tid == 0) {
            double gm = 0.0;
            for (int c = 0; c < K; ++c) gm = fmax(gm, w.cpart[WC_COUNT * (size_t)c + WC_GMAX]);
            for (int ch = 0; ch < D.n_chunks; ++ch) gm = fmax(gm, w.ppart[WP_COUNT * (size_t)ch + WP_GMAX]);
            if (first) {
                double cost = 0.0;
                for (int c = 0; c < K; ++c) cost += w.cpart[WC_COUNT * (size_t)c + WC_COST];
                st->x_cost = st->initial_cost = cost;
                st->rows = 1;
                if (!isfinite(cost)) lm_finish(st, BA_FAILURE, BA_REASON_NON_FINITE_COST, cost);
                if (o.g2o) {
                    // g2o's computeLambdaInit: tau times the largest diagonal entry of J'J
                    double hmax = 0.0;
                    for (int k = 0; k < n; ++k) hmax = fmax(hmax, w.udiag[k]);
                    for (int p = 0; p < P; ++p) {
                        const double* V = w.V + 6 * (size_t)p;
                        hmax = fmax(hmax, fmax(V[0], fmax(V[3], V[5])));
                    }
                    st->radius = 1.0 / (o.g2o_tau * hmax);  // radius = 1 / lambda
                }
            }
            st->gmax = gm;
            st->need_lin = 0;
            st->first_lin = 0;
            if (!st->done && first && st->iteration >= o.max_iterations)
                lm_finish(st, BA_NO_CONVERGENCE, BA_REASON_MAX_ITERATIONS, st->iteration);
            if (!o.g2o && !st->done && gm <= o.gtol) lm_finish(st, BA_CONVERGENCE, BA_REASON_GRADIENT_TOLERANCE, gm);
            if (!o.g2o && !st->done && st->radius <= o.min_radius)
                lm_finish(st, BA_CONVERGENCE, BA_REASON_MIN_RADIUS, st->radius);
        }
        if (tid == 0 && st->done) st->t_done = wall_clock64();
        __syncthreads();
        if (st->done) return;
    }
    if (tid == 0) bad = 0;
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
            bad = 1;
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
    if (tid == 0) {
        st->invalid = bad;
        st->have_step = 0;
    }
}

// Per (window, camera): records of the reduced system and b_c.
template <typename Real, int NT, int M>
__global__ void __launch_bounds__(NT) k_wrec_cam(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                 const int2* __restrict__ ctask, int ntask, Opts o) {
    constexpr int NW = NT / 64;
    __shared__ double lds[9 * NW];
    __shared__ double dacc[NW][45];  // per wave: upper triangle of sum J~_c'(I - G)J~_c
    __shared__ double red[NW][16][45];  // per wave: one round's terms summed over lane quads
    __shared__ CameraPrep<double> Psh;
    __shared__ StereoRig Gsh;
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const int2 tk = ctask[t];
    const WinDesc D = descs[tk.x];
    const WinPtrs w = win_ptrs(D, arena);
    const WinState* st = w.st;
    if (st->done || st->invalid || tk.y >= w.K) return;
    Real* js = at<Real>(arena, D.o_js);
    const int c = tk.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the camera's range and the first round's point index are requested before the barrier
    const int j_beg = w.cam_off[c], j_end = w.cam_off[c + 1];
    int pn = j_beg + (int)threadIdx.x < j_end ? w.co_pt[j_beg + threadIdx.x] : 0;
    if (threadIdx.x == 0) Psh = w.prep[c];
    win_load_rig<M>(w, Gsh);
    double scc[9];
    for (int i = 0; i < 9; ++i) scc[i] = w.sc[9 * (size_t)c + i];
    double bacc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) bacc[i] = 0.0;
    for (int k = threadIdx.x; k < NW * 45; k += NT) (&dacc[0][0])[k] = 0.0;
    __syncthreads();
    // rounds of NT observations, every lane taking part in the per-round butterflies
    for (int j0r = j_beg; j0r < j_end; j0r += NT) {
        const int j = j0r + threadIdx.x;
        const int p = pn;
        if (j + NT < j_end) pn = w.co_pt[j + NT];
        double Jd[18], g00 = 0.0, g01 = 0.0, g11 = 0.0;
#pragma unroll
        for (int i = 0; i < 18; ++i) Jd[i] = 0.0;
        if (j < j_end) {
            double r[2], Jc[18], Jp[6];
            win_scaled<M>(w, Psh, Gsh, j, p, o, scc, r, Jc, Jp);
            const double* Vi = w.Vinv + 6 * (size_t)p;
            const double j0[3] = {Jp[0], Jp[1], Jp[2]}, j1[3] = {Jp[3], Jp[4], Jp[5]};
            double h0[3], h1[3];
            sym3_mul<double>(Vi, j0, h0);
            sym3_mul<double>(Vi, j1, h1);
            Real* out = js + kWinJs * (size_t)j;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                out[2 * i] = (Real)Jc[i];
                out[2 * i + 1] = (Real)Jc[9 + i];
            }
            for (int i = 0; i < 3; ++i) {
                out[18 + i] = (Real)h0[i];
                out[21 + i] = (Real)h1[i];
            }
#pragma unroll
            for (int i = 0; i < 6; ++i) out[24 + i] = (Real)Jp[i];
            g00 = 1.0 - (j0[0] * h0[0] + j0[1] * h0[1] + j0[2] * h0[2]);
            g01 = -(j0[0] * h1[0] + j0[1] * h1[1] + j0[2] * h1[2]);
            g11 = 1.0 - (j1[0] * h1[0] + j1[1] * h1[1] + j1[2] * h1[2]);
            const double4 q = w.qp[p];
            const double e0 = r[0] - (Jp[0] * q.x + Jp[1] * q.y + Jp[2] * q.z);
            const double e1 = r[1] - (Jp[3] * q.x + Jp[4] * q.y + Jp[5] * q.z);
#pragma unroll
            for (int i = 0; i < 9; ++i) bacc[i] += Jc[i] * e0 + Jc[9 + i] * e1;
#pragma unroll
            for (int i = 0; i < 18; ++i) Jd[i] = Jc[i];
        }
        // the observation's own term of the camera's diagonal block: J~_c'(I - G)J~_c
        double t0[9], t1[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            t0[i] = g00 * Jd[i] + g01 * Jd[9 + i];
            t1[i] = g01 * Jd[i] + g11 * Jd[9 + i];
        }
        // the 45 terms summed over lane quads by two DPP steps, the 16 quad sums of each term then
        // added in order by lane e (an LDS transpose instead of 45 full-wave butterflies)
        int e = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int cc = r; cc < 9; ++cc, ++e) {
                double v = Jd[r] * t0[cc] + Jd[9 + r] * t1[cc];
                v += dpp_lane<kDppXor1>(v);
                v += dpp_lane<kDppXor2>(v);
                if ((lane & 3) == 0) red[wave][lane >> 2][e] = v;
            }
        wave_sync_lds();
        if (lane < 45) {
            double v = red[wave][0][lane];
#pragma unroll
            for (int m = 1; m < 16; ++m) v += red[wave][m][lane];
            dacc[wave][lane] += v;
        }
        wave_sync_lds();
    }
    wg_sum_fixed<9, NT>(bacc, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < 9; ++i) w.bc[9 * (size_t)c + i] = bacc[i];
    // the diagonal block of S: D_c^2 + sum over the camera's observations (the waves in order);
    // the pair kernel writes the off-diagonal blocks only
    if (threadIdx.x < 45) {
        const int k = threadIdx.x;
        int r = 0, rem = k;
        while (rem >= 9 - r) {
            rem -= 9 - r;
            ++r;
        }
        const int cc = r + rem;
        double v = dacc[0][k];
        for (int q = 1; q < NW; ++q) v += dacc[q][k];
        if (r == cc) v += w.dc2[9 * (size_t)c + r];
        double* S = at<double>(arena, D.o_S);
        const int n = 9 * w.K;
        S[(size_t)(9 * c + r) * n + 9 * c + cc] = v;
        S[(size_t)(9 * c + cc) * n + 9 * c + r] = v;
    }
}

// Per (window, block group): blocks first + stride t of the window's list.
template <typename Real, int NT>
__global__ void __launch_bounds__(NT) k_wschur(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                               const int4* __restrict__ stask, int ntask) {
    extern __shared__ __align__(32) unsigned char smem[];
    __shared__ int counter;
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const int4 tk = stask[t];
    const WinDesc D = descs[tk.x];
    const WinPtrs w = win_ptrs(D, arena);
    const WinState* st = w.st;
    if (st->done || st->invalid || tk.y >= D.n_blocks || tk.y + (int64_t)tk.z * (tk.w - 1) >= (int64_t)D.n_blocks)
        return;
    if (threadIdx.x == 0) counter = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    win_schur<Real>(w.blk, w.pairs, at<Real>(arena, D.o_js), w.dc2, at<double>(arena, D.o_S), 9 * w.K, tk.y, tk.z,
                          tk.w, &counter, reinterpret_cast<Real*>(smem + (size_t)wave * win_stage_bytes<Real>()));
}

// Per window: factor S (FP64) and solve for the camera step y_c (into yc).  The packed factor lives
// in LDS when the window's reduced system has at most lds_n unknowns and in the window's own global
// buffer (L2-resident at these sizes) otherwise; the arithmetic is the same either way.
template <int NT>
__global__ void __launch_bounds__(NT) k_wsolve(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                               const int* __restrict__ wtask, int ntask, int lds_n) {
    extern __shared__ __align__(32) unsigned char smem[];
    __shared__ double xv[kWinMaxN], dinv[kWinMaxN], xch[kWinNb];
    __shared__ int bad;
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const WinDesc D = descs[wtask[t]];
    const WinPtrs w = win_ptrs(D, arena);
    WinState* st = w.st;
    if (st->done || st->invalid) return;
    const int K = w.K, n = 9 * K, tid = threadIdx.x;
    const double* Sg = at<double>(arena, D.o_S);
    double* S = n > lds_n ? at<double>(arena, D.o_L) : reinterpret_cast<double*>(smem);
    const int ntri = n * (n + 1) / 2;
    // packed copy of the lower triangle, eight loads in flight per thread
    constexpr int U = 8;
    for (int b0 = tid; b0 < ntri; b0 += U * NT) {
        double v[U];
        int ix[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = min(b0 + u * NT, ntri - 1);
            int i = (int)((sqrtf(8.0f * idx + 1.0f) - 1.0f) * 0.5f);
            while (tri(i + 1) <= idx) ++i;
            while (tri(i) > idx) --i;
            ix[u] = idx;
            v[u] = Sg[(size_t)i * n + (idx - tri(i))];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b0 + u * NT < ntri) S[ix[u]] = v[u];
    }
    if (tid == 0) bad = 0;
    __syncthreads();
    win_cholesky<double, NT>(S, n, dinv, xch, &bad);
    if (bad) return;  // have_step stays 0: an invalid step
    for (int k = tid; k < n; k += NT) xv[k] = w.bc[k];
    __syncthreads();
    win_chol_solve<double, NT>(S, dinv, n, xv);
    for (int k = tid; k < n; k += NT) w.yc[k] = xv[k];
    if (tid == 0) st->have_step = 1;
}

// Per (window, camera): u = J~_p'(r~ - J~_c y_c) into slot order.
template <int NT, int M>
__global__ void __launch_bounds__(NT) k_wback_cam(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                  const int2* __restrict__ ctask, int ntask, Opts o) {
    __shared__ CameraPrep<double> Psh;
    __shared__ StereoRig Gsh;
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const int2 tk = ctask[t];
    const WinDesc D = descs[tk.x];
    const WinPtrs w = win_ptrs(D, arena);
    const WinState* st = w.st;
    if (st->done || !st->have_step || tk.y >= w.K) return;
    const int c = tk.y;
    const int j_end = w.cam_off[c + 1];
    int j = w.cam_off[c] + threadIdx.x;
    int pn = j < j_end ? w.co_pt[j] : 0;  // the next round's point index, one round ahead
    if (threadIdx.x == 0) Psh = w.prep[c];
    win_load_rig<M>(w, Gsh);
    double scc[9], y[9];
    for (int i = 0; i < 9; ++i) {
        scc[i] = w.sc[9 * (size_t)c + i];
        y[i] = w.yc[9 * (size_t)c + i];
    }
    __syncthreads();
    for (; j < j_end; j += NT) {
        const int p = pn;
        if (j + NT < j_end) pn = w.co_pt[j + NT];
        double r[2], Jc[18], Jp[6];
        win_scaled<M>(w, Psh, Gsh, j, p, o, scc, r, Jc, Jp);
        double a0 = r[0], a1 = r[1];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            a0 -= Jc[i] * y[i];
            a1 -= Jc[9 + i] * y[i];
        }
        pv_at(w, 0, w.slot_of[j]) =
            make_double4(Jp[0] * a0 + Jp[3] * a1, Jp[1] * a0 + Jp[4] * a1, Jp[2] * a0 + Jp[5] * a1, 0.0);
    }
}

// One wave per (window, chunk): y_p, candidate points, |dx_p|^2 and |x_p|^2 of the chunk.
__global__ void __launch_bounds__(64) k_wback_pt(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                 const int2* __restrict__ ptask, int ntask, Opts o) {
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const int2 tk = ptask[t];
    const WinDesc D = descs[tk.x];
    const WinPtrs w = win_ptrs(D, arena);
    const WinState* st = w.st;
    if (st->done || !st->have_step || tk.y >= D.n_chunks) return;
    const int ch = tk.y, p = 64 * ch + threadIdx.x;
    double dx2 = 0.0, x2 = 0.0, model = 0.0;
    if (p < w.P) {
        double tt[3], y[3];
        win_slot_sum3(w, p, tt);
        sym3_mul<double>(w.Vinv + 6 * (size_t)p, tt, y);
        const double4 x4 = w.pt_x[p];
        const double xs[3] = {x4.x, x4.y, x4.z};
        const double* V = w.V + 6 * (size_t)p;
        const double vd[3] = {V[0], V[3], V[5]};
        const double radius = st->radius;
        double cand[3];
        for (int i = 0; i < 3; ++i) {
            w.yp[3 * (size_t)p + i] = y[i];
            const double s = w.sp[3 * (size_t)p + i];
            const double dl = -s * y[i];
            cand[i] = xs[i] + dl;
            dx2 += dl * dl;
            x2 += xs[i] * xs[i];
            // the point's share of the model cost change y'(g~ + D^2 y) / 2 (D^2 as k_wdamp forms it)
            const double d2 = fmin(fmax(s * vd[i] * s, o.min_diag), o.max_diag) / radius;
            model += y[i] * (s * w.gp[3 * (size_t)p + i] + d2 * y[i]);
        }
        w.pt_c[p] = make_double4(cand[0], cand[1], cand[2], 0.0);
    }
    dx2 = wave_sum(dx2);
    x2 = wave_sum(x2);
    model = wave_sum(model);
    if (threadIdx.x == 0) {
        w.ppart[WP_COUNT * (size_t)ch + WP_DX2] = dx2;
        w.ppart[WP_COUNT * (size_t)ch + WP_X2] = x2;
        w.ppart[WP_COUNT * (size_t)ch + WP_MODEL] = 0.5 * model;
    }
}

// Per (window, camera): candidate camera, FP64 candidate cost of the camera's observations, the
// camera's share of the model cost change and |dx_c|^2, |x_c|^2.  The model cost change of the
// step, -(J~s)'(r~ + J~s/2) summed over the residuals (step s = -y in Jacobi-scaled coordinates), is
// formed without the Jacobian: the damped system (J~'J~ + D^2) y = g~ = J~'r~ is solved exactly
// (Cholesky of the reduced system, exact back-substitution), so it equals y'(g~ + D^2 y) / 2, a
// sum over parameters (cameras here, points in k_wback_pt).
template <int NT, int M>
__global__ void __launch_bounds__(NT) k_wcand_cam(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                  const int2* __restrict__ ctask, int ntask, Opts o) {
    __shared__ double lds[NT / 64];
    __shared__ CameraPrep<double> Psh;  // the candidate's prep
    __shared__ StereoRig Gsh;
    __shared__ double csh[9];
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const int2 tk = ctask[t];
    const WinDesc D = descs[tk.x];
    const WinPtrs w = win_ptrs(D, arena);
    const WinState* st = w.st;
    if (st->done || !st->have_step || tk.y >= w.K) return;
    const int c = tk.y;
    // the first round's point index and candidate point are requested before the barriers
    const int j_end = w.cam_off[c + 1];
    int j = w.cam_off[c] + threadIdx.x;
    double4 x4n = make_double4(0.0, 0.0, 0.0, 0.0);
    if (j < j_end) x4n = w.pt_c[w.co_pt[j]];
    if (threadIdx.x < 9) {
        const size_t k = 9 * (size_t)c + threadIdx.x;
        csh[threadIdx.x] = w.cam_x[k] - w.sc[k] * w.yc[k];
    }
    win_load_rig<M>(w, Gsh);
    __syncthreads();
    if (threadIdx.x == 0) {
        if constexpr (M == kModelStereo) {
            // pose <- exp(xi) pose, xi = (rho, phi) = the step's first six entries
            const double* x = w.cam_x + 9 * (size_t)c;
            double xi[6];
            for (int i = 0; i < 6; ++i) xi[i] = -w.sc[9 * (size_t)c + i] * w.yc[9 * (size_t)c + i];
            se3_left_update(xi, x, x + 3, csh, csh + 3);
        }
        for (int i = 0; i < 9; ++i) w.cam_c[9 * (size_t)c + i] = csh[i];
        camera_prep<double>(csh, Psh);
    }
    __syncthreads();
    double acc[1] = {0.0};  // candidate cost
    for (; j < j_end; j += NT) {
        const double4 x4 = x4n;
        if (j + NT < j_end) x4n = w.pt_c[w.co_pt[j + NT]];
        const double X[3] = {x4.x, x4.y, x4.z};
        double r[2];
        win_model_residual<M>(w, Psh, Gsh, j, X, r);
        acc[0] += half_rho(r[0] * r[0] + r[1] * r[1], o.loss, o.loss_a);
    }
    wg_sum_fixed<1, NT>(acc, lds);
    if (threadIdx.x == 0) {
        double dx2 = 0.0, x2 = 0.0, model = 0.0;
        for (int i = 0; i < 9; ++i) {
            const size_t k = 9 * (size_t)c + i;
            const double y = w.yc[k], dl = -w.sc[k] * y, x = w.cam_x[k];
            dx2 += dl * dl;
            x2 += x * x;
            model += y * (w.sc[k] * w.gc[k] + w.dc2[k] * y);
        }
        double* cp = w.cpart + WC_COUNT * (size_t)c;
        cp[WC_MODEL] = 0.5 * model;
        cp[WC_CAND] = acc[0];
        cp[WC_DX2] = dx2;
        cp[WC_X2] = x2;
    }
}

// The device wall clock at the start of a batch run (windows' finishing times are relative to it).
__global__ void k_wclock(unsigned long long* t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t = wall_clock64();
}

// The Ceres LM rule on a window's trial (k_decide's, without the trace): returns whether the
// candidate is accepted.
__device__ __forceinline__ int ceres_decide(WinState* st, const Opts& o, bool have_step, double model, double cand,
                                            double dx2, double x2) {
    int accepted = 0;
    const bool valid = have_step && model > 0.0 && isfinite(dx2);
    bool record = true;
    if (!valid) {
        if (++st->consecutive_invalid >= o.max_invalid) {
            lm_finish(st, BA_FAILURE, BA_REASON_INVALID_STEPS, st->consecutive_invalid);
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
            lm_finish(st, BA_CONVERGENCE, BA_REASON_PARAMETER_TOLERANCE, step_norm / (x_norm + o.ptol));
            record = false;
        } else {
            const double cost_change = st->x_cost - cand;
            if (fabs(cost_change) <= o.ftol * st->x_cost) {
                lm_finish(st, BA_CONVERGENCE, BA_REASON_FUNCTION_TOLERANCE, fabs(cost_change) / st->x_cost);
                record = false;
            } else {
                const double rho = cost_change / model;
                if (rho > o.min_rel_decrease) {
                    st->x_cost = cand;
                    const double tt = 2.0 * rho - 1.0;
                    st->radius = fmin(o.max_radius, st->radius / fmax(1.0 / 3.0, 1.0 - tt * tt * tt));
                    st->decrease_factor = 2.0;
                    st->need_lin = 1;
                    st->num_successful++;
                    accepted = 1;
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
        if (st->iteration >= o.max_iterations)
            lm_finish(st, BA_NO_CONVERGENCE, BA_REASON_MAX_ITERATIONS, st->iteration);
        else if (!accepted && st->gmax <= o.gtol)
            lm_finish(st, BA_CONVERGENCE, BA_REASON_GRADIENT_TOLERANCE, st->gmax);
        else if (!accepted && st->radius <= o.min_radius)
            lm_finish(st, BA_CONVERGENCE, BA_REASON_MIN_RADIUS, st->radius);
    }
    return accepted;
}

// g2o's OptimizationAlgorithmLevenberg on a window's trial (k_decide's g2o branch): rho = (chi2 -
// chi2_new) / (x'(lambda x + b) + 1e-3) with chi2 = 2 cost and x'(lambda x + b) = 2 model; accept
// when rho > 0, lambda *= max(1/3, min(2/3, 1 - (2 rho - 1)^3)), else lambda *= nu, nu *= 2; an
// iteration ends with an accepted trial, the solve after g2o_max_trials rejected ones (radius =
// 1 / lambda, decrease_factor = nu); every trial is a row.
__device__ __forceinline__ int g2o_decide(WinState* st, const Opts& o, bool have_step, double model, double cand) {
    int accepted = 0;
    if (!isfinite(cand)) cand = DBL_MAX;
    const double rho = have_step ? 2.0 * (st->x_cost - cand) / (2.0 * model + 1e-3) : -1.0;
    if (have_step && rho > 0.0 && cand < DBL_MAX) {
        const double tt = 2.0 * rho - 1.0;
        st->radius = st->radius / fmax(1.0 / 3.0, fmin(2.0 / 3.0, 1.0 - tt * tt * tt));
        st->decrease_factor = 2.0;
        st->x_cost = cand;
        st->need_lin = 1;
        st->g2o_iter++;
        st->g2o_trials = 0;
        st->num_successful++;
        accepted = 1;
    } else {
        st->radius = st->radius / st->decrease_factor;
        st->decrease_factor *= 2.0;
        st->g2o_trials++;
        st->num_unsuccessful++;
    }
    st->rows++;
    if (st->g2o_iter >= o.max_iterations)
        lm_finish(st, BA_NO_CONVERGENCE, BA_REASON_MAX_ITERATIONS, st->g2o_iter);
    else if (st->g2o_trials >= o.g2o_max_trials || rho == 0.0 || !isfinite(st->radius) || !(st->radius > 0.0))
        lm_finish(st, BA_CONVERGENCE, BA_REASON_G2O_STOP, rho);
    return accepted;
}

// Per window: the LM decision (Ceres or g2o rules) and, for an accepted step, the new parameters.
template <int NT>
__global__ void __launch_bounds__(NT) k_wdecide(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                const int* __restrict__ wtask, int ntask, Opts o) {
    __shared__ int accepted;
    const int t = blockIdx.x;
    if (t >= ntask) return;
    const WinDesc D = descs[wtask[t]];
    const WinPtrs w = win_ptrs(D, arena);
    WinState* st = w.st;
    if (st->done) return;
    if (threadIdx.x == 0) {
        accepted = 0;
        const bool have_step = st->have_step != 0;
        double model = 0.0, cand = 0.0, dx2 = 0.0, x2 = 0.0;
        if (have_step) {
            for (int c = 0; c < w.K; ++c) {
                const double* cp = w.cpart + WC_COUNT * (size_t)c;
                model += cp[WC_MODEL];
                cand += cp[WC_CAND];
                dx2 += cp[WC_DX2];
                x2 += cp[WC_X2];
            }
            for (int ch = 0; ch < D.n_chunks; ++ch) {
                dx2 += w.ppart[WP_COUNT * (size_t)ch + WP_DX2];
                x2 += w.ppart[WP_COUNT * (size_t)ch + WP_X2];
                model += w.ppart[WP_COUNT * (size_t)ch + WP_MODEL];
            }
        }
        st->iteration++;
        if (o.g2o)
            accepted = g2o_decide(st, o, have_step, model, cand);
        else
            accepted = ceres_decide(st, o, have_step, model, cand, dx2, x2);
        if (st->done) st->t_done = wall_clock64();
    }
    __syncthreads();
    if (accepted) {
        const int n = 9 * w.K;
        for (int k = threadIdx.x; k < n; k += NT) w.cam_x[k] = w.cam_c[k];
        for (int p = threadIdx.x; p < w.P; p += NT) w.pt_x[p] = w.pt_c[p];
    }
}

}  // namespace mba
