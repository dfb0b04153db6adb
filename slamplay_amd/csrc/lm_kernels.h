// lm_kernels.h -- device kernels of the Levenberg-Marquardt bundle-adjustment solve.
//
// Device layout (built once per problem by the host, see solver.hip):
//   * observations in camera order (CSR by camera, cam_off): co_pt (point), co_uv (x,y);
//     every Jacobian is recomputed in camera-order kernels, where the camera -- and so its
//     CameraPrep -- is uniform across the workgroup, from 12-20 bytes of observation data plus a
//     gathered point.  No per-observation Jacobian is stored.
//   * points in chunks of 64 (one wave).  Slot chunk_off[ch] + 64 r + l belongs to the r-th
//     observation of point 64 ch + l; slot_of[j] maps a camera-order observation to its slot.
//     Camera kernels write their per-observation partials for the points straight into slot
//     order, so point kernels read them as coalesced streams with no indirection (padding slots
//     stay zero).  Chunks hold points of similar track length, which keeps the padding small.
//   * internal point order: by first observing camera, then track length (longest first), so a
//     camera's observations touch clustered points and a wave's 64 points have similar tracks;
//   * parameters FP64: cameras SoA cam_x[i * nc + c]; points AoS4 pt_x[4 p + i] (one aligned
//     32-byte record per gather); per-point and per-observation working vectors are 4-wide too.
//   * workgroups are remapped so that each XCD (own L2) walks a contiguous range of cameras or
//     point chunks (xcd_remap).
// All reductions are fixed-order tree reductions in registers/LDS: no atomics, and results are
// bitwise reproducible run to run.
//
// The LM control (Ceres TrustRegionMinimizer + LevenbergMarquardtStrategy semantics as used by
// test_bundle_adjustment_ceres.cpp:96-100) runs in k_decide on the device; every other kernel reads
// its flags and exits early when it has nothing to do, so a fixed launch sequence per iteration
// can be enqueued (or captured) without host round trips.
#pragma once

#include <cfloat>

#include "common.h"
#include "snavely.h"
#include "stereo.h"

namespace mba {

constexpr int kWG = 256;    // threads per workgroup for every solver kernel
constexpr int kWaves = kWG / 64;
constexpr int kXcds = 8;    // MI355X: 8 XCDs, workgroups dispatched round-robin over them

// Block b runs on XCD b % 8; give XCD k the contiguous index range [k per, (k + 1) per) so that
// neighbouring cameras (which share points) and neighbouring point chunks share one L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int per = (n + kXcds - 1) / kXcds;
    return (b % kXcds) * per + b / kXcds;
}
inline int xcd_grid(int n) { return kXcds * ((n + kXcds - 1) / kXcds); }

// 4-wide aligned record of per-point / per-observation working vectors (w unused).
template <typename Real>
struct alignas(4 * sizeof(Real)) R4 {
    Real x, y, z, w;
};

// Layout of the per-workgroup partial-sum slots (partial[slot * stride + block]).
enum PartialSlot {
    PS_GMAX = 0,   // max |gradient|
    PS_COST = 1,   // cost
    PS_MODEL = 2,  // model cost change
    PS_DX2 = 3,    // |delta|^2
    PS_X2 = 4,     // |x|^2
    PS_A = 5,      // PCG scalar partial A (rz / pUp)
    PS_B = 6,      // PCG scalar partial B (rr / tz)
    PS_COUNT = 7
};

struct LMState {
    double radius, decrease_factor;
    double x_cost, cand_cost, model_change;
    double gmax;
    double step_norm, x_norm;
    double pcg_bb;
    int iteration, consecutive_invalid, done, termination;
    int need_lin, first_lin, accepted, step_invalid;
    int num_successful, num_unsuccessful;
    int pcg_iters, pcg_done, total_pcg_iters, trace_len;
    int g2o_iter, g2o_trials;  // g2o mode: outer iterations done, rejected trials of the current one
    int reason;                // ba_termination_reason of the test that ended the solve
    double reason_value;       // the quantity that test compared
    unsigned long long t0;     // device wall clock at the first evaluation (trace row times)
};

// Ends the solve with `termination`, recording which test fired and the value it tested.
template <class State>
__device__ __forceinline__ void lm_finish(State* st, int termination, int reason, double value) {
    st->done = 1;
    st->termination = termination;
    st->reason = reason;
    st->reason_value = value;
}

// Device wall-clock ticks since the first evaluation (converted to seconds on the host).
__device__ __forceinline__ double lm_ticks(const LMState* st) { return (double)(wall_clock64() - st->t0); }

struct Opts {
    int loss;
    int max_iterations;
    int max_invalid;
    int jacobi_scaling;
    int pcg_max_iterations;
    double loss_a;
    double ftol, gtol, ptol;
    double max_radius, min_radius, min_rel_decrease;
    double min_diag, max_diag;
    double pcg_tol;
    int g2o;             // the g2o driver's LM (BA_MODE_G2O)
    int g2o_max_trials;  // rejected trials per iteration before the solve stops
    double g2o_tau;      // initial lambda = tau * max diag(J'J)
};

template <typename Real>
struct Dev {
    int nc, np, no, n_chunks, n_pt_blocks;
    Opts o;
    const int* cam_off;
    const int* co_pt;
    const double* co_uv64;
    const Real* co_uv;
    R4<Real>* co_rec;          // per observation (camera order): X - c (point relative to the camera centre)
                               // and the Huber weight sqrt(rho'), at the linearization (direction-form PCG)
    const int* chunk_off;
    const int* chunk_len;
    const int* slot_of;
    double *cam_x, *pt_x, *cam_c, *pt_c, *dcam, *dpt;
    CameraPrep<Real>* prep;
    CameraPrep<double>* prep_c;
    double *U, *gc, *sc, *dc2; // per camera: diag(Jc'Jc), Jc'r, Jacobi scale, damping D_c^2 (9 each)
    double *M, *Minv;          // per camera: 81 each (Schur-Jacobi block and its inverse)
    double *bc, *px, *pr, *pz, *pp, *qc;  // per camera: 9 each (PCG vectors, FP64)
    double *V, *gp, *sp, *dp2; // per point: 6, 3, 3, 3
    double* Vinv64;            // per point: 6 (FP64 inverse of the damped block)
    Real* VinvR;               // per point: 6 (same, in working precision)
    R4<Real>* qp;              // per point: Vinv S g_p
    double4* zs;               // per point: S_p z_p of the current PCG direction (FP64 in both precisions)
    double4* pv;               // per slot, 3 records: J_p'J_p (sym 6), J_p'r (3), padding
    R4<Real>* u;               // per slot: J_p'(J_c S_c v_c)
    double* cam_part;          // PS_COUNT * nc
    double* pt_part;           // PS_COUNT * n_pt_blocks
    double* upd_part;          // 2 * n_upd: PCG r.z and r.r per block of kCamsPerUpd cameras
    double* rz_hist;           // r.z at the start of each PCG iteration
    int n_upd;
    LMState* st;
    ba_iteration* trace;
    int trace_cap;
    int model;                 // kModelSnavely / kModelSnavelyG2o / kModelStereo (below)
    const int* co_side;        // stereo: camera side (0 left, 1 right) per camera-order observation
    StereoRig rig;             // stereo
};

// Residual models, a compile-time parameter of the kernels that evaluate residuals or update the
// cameras, so each solver runs code with only its own model's registers live.
constexpr int kModelSnavely = 0;     // Snavely camera, additive update (Ceres)
constexpr int kModelSnavelyG2o = 1;  // Snavely camera, exp(delta) R rotation update (g2o driver)
constexpr int kModelStereo = 2;      // stereo rig, SE3 pose update (full_slam)

// The instantiation of a kernel for residual model m.
template <typename K>
inline K pick_model(int m, K snavely, K snavely_g2o, K stereo) {
    return m == kModelStereo ? stereo : (m == kModelSnavelyG2o ? snavely_g2o : snavely);
}

// ---------------------------------------------------------------------------------------------
// Reductions

// Value of v in lane `lane` (a wave-uniform index) without a trip through LDS.
__device__ __forceinline__ double read_lane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float read_lane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// 1 / sqrt(d) for d > 0: the hardware estimate and Newton steps r += r (1 - d r^2) / 2 (two in
// FP64, whose estimate has about 23 correct bits; one in FP32).
__device__ __forceinline__ double inv_sqrt(double d) {
    double r = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int i = 0; i < 2; ++i) r = fma(0.5 * r, fma(-d * r, r, 1.0), r);
    return r;
}
__device__ __forceinline__ float inv_sqrt(float d) {
    const float r = __builtin_amdgcn_rsqf(d);
    return fmaf(0.5f * r, fmaf(-d * r, r, 1.0f), r);
}

// Orders the LDS accesses of one wave around an exchange through LDS (the
// wave's LDS operations complete in order; the fences keep the compiler from moving them).
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave reductions on the VALU: four butterfly steps inside each row of 16 lanes by DPP lane
// permutations (partners lane ^ 1, lane ^ 2, the mirrored lane of the half-row and of the row), so
// every lane of a row holds the row's total, then the four row totals read with v_readlane and added
// in a fixed order.  (Shuffles through ds_bpermute cost an LDS round trip per step; a workgroup
// reducing a few dozen values spent most of its time waiting on them.)  Every lane of the wave must
// be active.
constexpr int kDppXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppMirror = 0x140;     // row_mirror

template <int CTRL>
__device__ __forceinline__ double dpp_lane(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over each row of 16 lanes; every lane of the row gets its row's total.
template <typename T>
__device__ __forceinline__ T row16_sum(T v) {
    v += dpp_lane<kDppXor1>(v);
    v += dpp_lane<kDppXor2>(v);
    v += dpp_lane<kDppHalfMirror>(v);
    v += dpp_lane<kDppMirror>(v);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    v = row16_sum(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Sum over each half-wave of 32 lanes (lanes 0-31 and 32-63 reduced separately).
template <typename T>
__device__ __forceinline__ T half_wave_sum(T v) {
    v = row16_sum(v);
    const T a = read_lane(v, 0) + read_lane(v, 16), b = read_lane(v, 32) + read_lane(v, 48);
    return (threadIdx.x & 32) ? b : a;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    T o = dpp_lane<kDppXor1>(v);
    v = o > v ? o : v;
    o = dpp_lane<kDppXor2>(v);
    v = o > v ? o : v;
    o = dpp_lane<kDppHalfMirror>(v);
    v = o > v ? o : v;
    o = dpp_lane<kDppMirror>(v);
    v = o > v ? o : v;
    const T a = read_lane(v, 0), b = read_lane(v, 16), c = read_lane(v, 32), e = read_lane(v, 48);
    const T ab = a > b ? a : b, ce = c > e ? c : e;
    return ab > ce ? ab : ce;
}

// sum_{k = threadIdx.x, + kWG, + 2 kWG, ... < n} a[k], added in that order, four loads in flight
// (clamped indices and a select instead of a branch around each load, so a load does not wait for
// the one before it).
__device__ __forceinline__ double strided_sum(const double* __restrict__ a, int n) {
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 4 * kWG) {
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = k + u * kWG;
            const double v = a[min(kk, n - 1)];
            x[u] = kk < n ? v : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) s += x[u];
    }
    return s;
}

// Sums N per-thread values over a workgroup of NW waves (256 threads by default). On return
// lds[i * NW] holds the total of value i for every thread (a barrier is included).
template <int N, typename T, int NW = kWaves>
__device__ __forceinline__ void block_sum(T (&v)[N], T* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const T s = wave_sum(v[i]);
        if (lane == 0) lds[i * NW + wave] = s;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += 64 * NW) {
        T s = lds[i * NW];
        for (int w = 1; w < NW; ++w) s += lds[i * NW + w];
        lds[i * NW] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ double block_sum1(double v, double* lds) {
    double a[1] = {v};
    block_sum<1>(a, lds);
    return lds[0];
}

__device__ __forceinline__ double block_max1(double v, double* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_max(v);
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double m = lds[0];
    for (int w = 1; w < kWaves; ++w) m = lds[w] > m ? lds[w] : m;
    __syncthreads();
    return m;
}

// Index of (i, k), i <= k, in a packed upper triangle of a 9x9 matrix (45 entries).
__host__ __device__ __forceinline__ constexpr int sym9(int i, int k) { return i * 9 - (i * (i - 1)) / 2 + (k - i); }

// ---------------------------------------------------------------------------------------------
// Small dense FP64 algebra (per-thread)

// In-place Cholesky of an n x n row-major SPD matrix. Returns false if not positive definite.
template <int N>
__device__ __forceinline__ bool chol_inplace(double (&A)[N * N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double d = A[j * N + j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= A[j * N + k] * A[j * N + k];
        if (!(d > 0.0) || !isfinite(d)) return false;
        d = sqrt(d);
        A[j * N + j] = d;
        const double inv = 1.0 / d;
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            double s = A[i * N + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= A[i * N + k] * A[j * N + k];
            A[i * N + j] = s * inv;
        }
    }
    return true;
}

// Inverse of an SPD matrix from its lower Cholesky factor L (A = L L'): A^-1 = L^-T L^-1.
template <int N>
__device__ __forceinline__ void chol_inverse(const double (&L)[N * N], double (&Ai)[N * N]) {
    double Li[N * N];  // L^-1 (lower)
#pragma unroll
    for (int j = 0; j < N; ++j) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (i < j) {
                Li[i * N + j] = 0.0;
            } else if (i == j) {
                Li[i * N + j] = 1.0 / L[i * N + i];
            } else {
                double s = 0.0;
#pragma unroll
                for (int k = j; k < i; ++k) s -= L[i * N + k] * Li[k * N + j];
                Li[i * N + j] = s / L[i * N + i];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = i; k < N; ++k) s += Li[k * N + i] * Li[k * N + j];
            Ai[i * N + j] = s;
            Ai[j * N + i] = s;
        }
}

// ---------------------------------------------------------------------------------------------
// Observation helpers

template <typename Real>
__device__ __forceinline__ void load_point(const double* __restrict__ pts, int p, Real X[3]) {
    const double4 v = reinterpret_cast<const double4*>(pts)[p];
    X[0] = (Real)v.x;
    X[1] = (Real)v.y;
    X[2] = (Real)v.z;
}

// Inputs of one observation (point index and position, measurement), loaded ahead of their use so
// that the index -> point gather of the next observation overlaps the current one's arithmetic.
template <typename Real>
struct ObsIn {
    int p, side;
    double X[3];
    Real u, v;
};

template <typename Real, int M>
__device__ __forceinline__ void load_obs(const Dev<Real>& d, int j, ObsIn<Real>& in) {
    in.p = d.co_pt[j];
    in.side = M == kModelStereo ? d.co_side[j] : 0;
    load_point<double>(d.pt_x, in.p, in.X);
    in.u = d.co_uv[2 * (size_t)j];
    in.v = d.co_uv[2 * (size_t)j + 1];
}

// Corrected residual and Jacobians from prefetched inputs (same arithmetic as corrected_jacobian).
template <typename Real, int M>
__device__ __forceinline__ void corrected_from(const Dev<Real>& d, const CameraPrep<Real>& P, const ObsIn<Real>& in,
                                               Real r[2], Real Jc[18], Real Jp[6]) {
    if constexpr (M == kModelStereo)
        stereo_residual_jacobian<Real>(P, in.X, d.rig, in.side, in.u, in.v, r, Jc, Jp);
    else
        snavely_residual_jacobian<Real>(P, in.X, in.u, in.v, r, Jc, Jp, M == kModelSnavelyG2o);
    const Real w = huber_sqrt_rho1<Real>(r[0] * r[0] + r[1] * r[1], d.o.loss, (Real)d.o.loss_a);
    r[0] *= w;
    r[1] *= w;
#pragma unroll
    for (int i = 0; i < 18; ++i) Jc[i] *= w;
#pragma unroll
    for (int i = 0; i < 6; ++i) Jp[i] *= w;
}

// Corrected (sqrt(rho')-weighted) residual and Jacobians of observation j of camera prep P (and
// the weight itself in *w_out when given).
template <typename Real, int M>
__device__ __forceinline__ void corrected_jacobian(const Dev<Real>& d, const CameraPrep<Real>& P, int j, int& p,
                                                   Real r[2], Real Jc[18], Real Jp[6], Real* w_out = nullptr) {
    p = d.co_pt[j];
    double X[3];
    load_point<double>(d.pt_x, p, X);
    if constexpr (M == kModelStereo)
        stereo_residual_jacobian<Real>(P, X, d.rig, d.co_side[j], d.co_uv[2 * (size_t)j], d.co_uv[2 * (size_t)j + 1], r, Jc, Jp);
    else
        snavely_residual_jacobian<Real>(P, X, d.co_uv[2 * (size_t)j], d.co_uv[2 * (size_t)j + 1], r, Jc, Jp,
                                        M == kModelSnavelyG2o);
    const Real w = huber_sqrt_rho1<Real>(r[0] * r[0] + r[1] * r[1], d.o.loss, (Real)d.o.loss_a);
    if (w_out) *w_out = w;
    r[0] *= w;
    r[1] *= w;
#pragma unroll
    for (int i = 0; i < 18; ++i) Jc[i] *= w;
#pragma unroll
    for (int i = 0; i < 6; ++i) Jp[i] *= w;
}

// ---------------------------------------------------------------------------------------------
// Linearization

template <typename Real>
__global__ void __launch_bounds__(64) k_cam_prep(Dev<Real> d) {
    if (!d.st->need_lin) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.nc) return;
    double cam[9];
    for (int i = 0; i < 9; ++i) cam[i] = d.cam_x[(size_t)i * d.nc + c];
    CameraPrep<Real> P;
    camera_prep<Real>(cam, P);
    d.prep[c] = P;
}

// One workgroup per camera: diag(Jc'Jc) and g_c = Jc'r (unscaled, corrected) -- the column norms
// for the Jacobi scaling and the LM diagonal -- plus each observation's point contributions
// Jp'Jp and Jp'r for k_lin_pt.
template <typename Real, int M>
__global__ void __launch_bounds__(kWG) k_lin_cam(Dev<Real> d) {
    if (!d.st->need_lin) return;
    __shared__ double lds[18 * kWaves];
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    const CameraPrep<Real> P = d.prep[c];
    double cc[3];  // camera centre
    camera_centre<Real>(P, cc);
    double acc[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) acc[i] = 0.0;
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6], w;
        corrected_jacobian<Real, M>(d, P, j, p, r, Jc, Jp, &w);
        if (M == kModelSnavely) {  // the direction-form PCG products' record
            const double4 X = reinterpret_cast<const double4*>(d.pt_x)[p];
            R4<Real> rec;
            rec.x = (Real)(X.x - cc[0]);
            rec.y = (Real)(X.y - cc[1]);
            rec.z = (Real)(X.z - cc[2]);
            rec.w = w;
            d.co_rec[j] = rec;
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            acc[i] += (double)(Jc[i] * Jc[i] + Jc[9 + i] * Jc[9 + i]);
            acc[9 + i] += (double)(Jc[i] * r[0] + Jc[9 + i] * r[1]);
        }
        // The point block's per-observation products are formed in FP64 from the (FP32 or FP64)
        // Jacobian: products of FP32 values are exact in FP64, so each contribution is exactly
        // positive semidefinite and V stays positive definite for nearly depth-free points.
        const double q[6] = {Jp[0], Jp[1], Jp[2], Jp[3], Jp[4], Jp[5]};
        const double r0 = r[0], r1 = r[1];
        double4* pv = d.pv + 3 * (size_t)d.slot_of[j];
        pv[0] = make_double4(q[0] * q[0] + q[3] * q[3], q[0] * q[1] + q[3] * q[4], q[0] * q[2] + q[3] * q[5],
                             q[1] * q[1] + q[4] * q[4]);
        pv[1] = make_double4(q[1] * q[2] + q[4] * q[5], q[2] * q[2] + q[5] * q[5], q[0] * r0 + q[3] * r1,
                             q[1] * r0 + q[4] * r1);
        pv[2] = make_double4(q[2] * r0 + q[5] * r1, 0.0, 0.0, 0.0);
    }
    block_sum<18>(acc, lds);
    const int t = threadIdx.x;
    if (t < 9) {
        d.U[9 * (size_t)c + t] = lds[t * kWaves];
        d.gc[9 * (size_t)c + t] = lds[(9 + t) * kWaves];
    }
    if (t == 0) {
        double gm = 0.0;
        for (int i = 0; i < 9; ++i) gm = fmax(gm, fabs(lds[(9 + i) * kWaves]));
        d.cam_part[PS_GMAX * (size_t)d.nc + c] = gm;
        if (d.st->first_lin)
            for (int i = 0; i < 9; ++i)
                d.sc[9 * (size_t)c + i] = d.o.jacobi_scaling ? 1.0 / (1.0 + sqrt(lds[i * kWaves])) : 1.0;
    }
}

// Point kernels: one workgroup per SELL chunk (64 points).  The 4 waves split the chunk's rows
// (wave w takes rows w, w + 4, ...; two rows in flight per lane) and combine through LDS, so a
// chunk's slots are streamed with 4x the memory parallelism of one wave walking all rows.  On
// return, lane l of wave 0 holds the sums of point 64 ch + l; other waves' sums are undefined.
template <int N, typename Load>
__device__ __forceinline__ void chunk_sum(int len, size_t base, double (&acc)[N], Load load, double* lds) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = 0.0;
    int r = wave;
    for (; r + kWaves < len; r += 2 * kWaves) {
        load(base + 64 * (size_t)r + lane, acc);
        load(base + 64 * (size_t)(r + kWaves) + lane, acc);
    }
    if (r < len) load(base + 64 * (size_t)r + lane, acc);
    if (wave > 0)
#pragma unroll
        for (int i = 0; i < N; ++i) lds[((wave - 1) * N + i) * 64 + lane] = acc[i];
    __syncthreads();
    if (wave == 0)
#pragma unroll
        for (int i = 0; i < N; ++i)
            for (int w = 0; w < kWaves - 1; ++w) acc[i] += lds[(w * N + i) * 64 + lane];
}

// chunk_sum over a stream of R4 records (x, y, z added), four rows in flight per lane: the rows of
// a wave are added in the same order as chunk_sum adds them.
template <typename Real>
__device__ __forceinline__ void chunk_sum_r4(int len, size_t base, const R4<Real>* __restrict__ src, double (&acc)[3],
                                             double* lds) {
    constexpr int U = 4;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    acc[0] = acc[1] = acc[2] = 0.0;
    // every trip requests U rows at once (rows past the chunk re-read its last row and add zero), so
    // a wave with fewer than U rows left still waits on memory once
    for (int r = wave; r < len; r += U * kWaves) {
        R4<Real> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[base + 64 * (size_t)min(r + u * kWaves, len - 1) + lane];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = r + u * kWaves < len;
            acc[0] += in ? (double)v[u].x : 0.0;
            acc[1] += in ? (double)v[u].y : 0.0;
            acc[2] += in ? (double)v[u].z : 0.0;
        }
    }
    if (wave > 0)
#pragma unroll
        for (int i = 0; i < 3; ++i) lds[((wave - 1) * 3 + i) * 64 + lane] = acc[i];
    __syncthreads();
    if (wave == 0)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            for (int w = 0; w < kWaves - 1; ++w) acc[i] += lds[(w * 3 + i) * 64 + lane];
}

// One workgroup per chunk: V_p = sum Jp'Jp, g_p = sum Jp'r per point.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_lin_pt(Dev<Real> d) {
    if (!d.st->need_lin) return;
    __shared__ double lds[(kWaves - 1) * 9 * 64];
    const int ch = xcd_remap(blockIdx.x, d.n_chunks);
    if (ch >= d.n_chunks) return;
    double a[9];
    chunk_sum<9>(d.chunk_len[ch], (size_t)d.chunk_off[ch], a,
                 [&](size_t slot, double (&acc)[9]) {
                     const double4* pv = d.pv + 3 * slot;
                     const double4 x = pv[0], y = pv[1], z = pv[2];
                     acc[0] += x.x;
                     acc[1] += x.y;
                     acc[2] += x.z;
                     acc[3] += x.w;
                     acc[4] += y.x;
                     acc[5] += y.y;
                     acc[6] += y.z;
                     acc[7] += y.w;
                     acc[8] += z.x;
                 },
                 lds);
    if (threadIdx.x >= 64) return;
    const int p = 64 * ch + threadIdx.x;
    double gm = 0.0;
    if (p < d.np) {
        for (int i = 0; i < 6; ++i) d.V[6 * (size_t)p + i] = a[i];
        for (int i = 0; i < 3; ++i) d.gp[3 * (size_t)p + i] = a[6 + i];
        gm = fmax(fabs(a[6]), fmax(fabs(a[7]), fabs(a[8])));
        if (d.st->first_lin) {
            const double dg[3] = {a[0], a[3], a[5]};
            for (int i = 0; i < 3; ++i) d.sp[3 * (size_t)p + i] = d.o.jacobi_scaling ? 1.0 / (1.0 + sqrt(dg[i])) : 1.0;
        }
    }
    gm = wave_max(gm);
    if (threadIdx.x == 0) d.pt_part[PS_GMAX * (size_t)d.n_pt_blocks + ch] = gm;
}

// FP64 cost 1/2 sum rho at the parameters whose prep is prep_c and points pts (one WG per camera).
template <typename Real, int M>
__global__ void __launch_bounds__(kWG) k_cost_cam(Dev<Real> d, const double* __restrict__ pts) {
    __shared__ double lds[kWaves];
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    const CameraPrep<double> P = d.prep_c[c];
    double acc = 0.0;
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        const int p = d.co_pt[j];
        double X[3];
        load_point<double>(pts, p, X);
        double r[2];
        if constexpr (M == kModelStereo)
            stereo_residual<double>(P, X, d.rig, d.co_side[j], d.co_uv64[2 * (size_t)j], d.co_uv64[2 * (size_t)j + 1], r);
        else
            snavely_residual<double>(P, X, d.co_uv64[2 * (size_t)j], d.co_uv64[2 * (size_t)j + 1], r);
        acc += half_rho(r[0] * r[0] + r[1] * r[1], d.o.loss, d.o.loss_a);
    }
    acc = block_sum1(acc, lds);
    if (threadIdx.x == 0) d.cam_part[PS_COST * (size_t)d.nc + c] = acc;
}

template <typename Real>
__global__ void __launch_bounds__(64) k_prep_from(Dev<Real> d, const double* __restrict__ cams) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.nc) return;
    double cam[9];
    for (int i = 0; i < 9; ++i) cam[i] = cams[(size_t)i * d.nc + c];
    CameraPrep<double> P;
    camera_prep<double>(cam, P);
    d.prep_c[c] = P;
}

// Reduces |g|max after a linearization, completes the trace row of the step that produced it and
// applies the gradient-tolerance test.  For the first linearization it also records iteration 0.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_post_lin(Dev<Real> d) {
    __shared__ double lds[kWaves];
    LMState* st = d.st;
    if (!st->need_lin) return;
    double gm = 0.0, cost = 0.0;
    for (int i = threadIdx.x; i < d.nc; i += kWG) {
        gm = fmax(gm, d.cam_part[PS_GMAX * (size_t)d.nc + i]);
        cost += d.cam_part[PS_COST * (size_t)d.nc + i];
    }
    for (int i = threadIdx.x; i < d.n_pt_blocks; i += kWG) gm = fmax(gm, d.pt_part[PS_GMAX * (size_t)d.n_pt_blocks + i]);
    gm = block_max1(gm, lds);
    cost = block_sum1(cost, lds);
    double hmax = 0.0;
    if (d.o.g2o && st->first_lin) {
        // g2o's computeLambdaInit: tau times the largest diagonal entry of J'J (cameras and points)
        __syncthreads();  // lds is reused
        for (int i = threadIdx.x; i < 9 * d.nc; i += kWG) hmax = fmax(hmax, d.U[i]);
        for (int p = threadIdx.x; p < d.np; p += kWG) {
            const double* V = d.V + 6 * (size_t)p;
            hmax = fmax(hmax, fmax(V[0], fmax(V[3], V[5])));
        }
        hmax = block_max1(hmax, lds);
    }
    if (threadIdx.x != 0) return;
    st->gmax = gm;
    if (st->first_lin) {
        if (d.o.g2o) st->radius = 1.0 / (d.o.g2o_tau * hmax);  // radius = 1 / lambda
        st->x_cost = cost;
        st->t0 = wall_clock64();
        if (st->trace_len < d.trace_cap) {
            ba_iteration row{};
            row.iteration = 0;
            row.step_is_successful = 1;
            row.step_is_valid = 1;
            row.cost = cost;
            row.gradient_max_norm = gm;
            row.trust_region_radius = st->radius;
            d.trace[st->trace_len] = row;
        }
        st->trace_len++;
        if (!isfinite(cost))
            lm_finish(st, BA_FAILURE, BA_REASON_NON_FINITE_COST, cost);
        else if (st->iteration >= d.o.max_iterations)
            lm_finish(st, BA_NO_CONVERGENCE, BA_REASON_MAX_ITERATIONS, st->iteration);
    } else if (st->trace_len > 0 && st->trace_len <= d.trace_cap) {
        d.trace[st->trace_len - 1].gradient_max_norm = gm;
    }
    if (!d.o.g2o && !st->done && gm <= d.o.gtol) lm_finish(st, BA_CONVERGENCE, BA_REASON_GRADIENT_TOLERANCE, gm);
    if (!d.o.g2o && !st->done && st->radius <= d.o.min_radius)
        lm_finish(st, BA_CONVERGENCE, BA_REASON_MIN_RADIUS, st->radius);
    st->need_lin = 0;
    st->first_lin = 0;
}

// ---------------------------------------------------------------------------------------------
// Damped blocks and the reduced camera system
//
// Every Schur quantity is evaluated in per-observation residual form.  With J~c = Jc S_c,
// J~p = Jp S_p (Jacobi-scaled, Huber-corrected), V*_p = sum J~p'J~p + D_p^2 and
// G_k = J~p V*^-1 J~p' (2x2) for observation k:
//   M_c   = D_c^2 + sum_k J~c'(I - G_k) J~c                      (Schur-Jacobi block)
//   b_c   = sum_k J~c'(r~_k - J~p q_p),   q_p = V*^-1 g~_p           (reduced gradient)
//   S v   = D_c^2 v_c + sum_k J~c' e_k,   e_k = J~c v_c - J~p z_p,  z = V*^-1 W~'v
//   v'Sv  = sum_k |e_k|^2 + sum_p z_p'D_p^2 z_p + sum_c v_c'D_c^2 v_c
//   y_p   = V*^-1 sum_k J~p'(r~_k - J~c y_c)                         (back-substitution)
// These equal U* - W V^-1 W', g - W V^-1 g_p, ... exactly, but never subtract two large
// accumulated sums: the cancellation happens inside one observation's 2-vector, so the FP32
// path stays accurate when the damping is small (large trust radius) and v'Sv cannot go
// negative through rounding.

// One thread per point: D_p^2 = clamp(diag(S V S))/radius, Vinv = (S V S + D_p^2)^-1 in FP64,
// q_p = Vinv (S g_p).
template <typename Real>
__global__ void __launch_bounds__(kWG) k_damp_pt(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done) return;
    const int p = blockIdx.x * kWG + threadIdx.x;
    if (p >= d.np) return;
    const double radius = st->radius;
    const double* V = d.V + 6 * (size_t)p;
    const double s[3] = {d.sp[3 * (size_t)p], d.sp[3 * (size_t)p + 1], d.sp[3 * (size_t)p + 2]};
    double A[9];
    A[0] = s[0] * V[0] * s[0];
    A[1] = A[3] = s[0] * V[1] * s[1];
    A[2] = A[6] = s[0] * V[2] * s[2];
    A[4] = s[1] * V[3] * s[1];
    A[5] = A[7] = s[1] * V[4] * s[2];
    A[8] = s[2] * V[5] * s[2];
    for (int i = 0; i < 3; ++i) {
        const double dd = fmin(fmax(A[4 * i], d.o.min_diag), d.o.max_diag) / radius;
        d.dp2[3 * (size_t)p + i] = dd;
        A[4 * i] += dd;
    }
    if (!chol_inplace<3>(A)) {
        st->step_invalid = 1;
        return;
    }
    double Ai[9];
    chol_inverse<3>(A, Ai);
    const double sym[6] = {Ai[0], Ai[1], Ai[2], Ai[4], Ai[5], Ai[8]};
    for (int i = 0; i < 6; ++i) {
        d.Vinv64[6 * (size_t)p + i] = sym[i];
        d.VinvR[6 * (size_t)p + i] = (Real)sym[i];
    }
    const double g[3] = {s[0] * d.gp[3 * (size_t)p], s[1] * d.gp[3 * (size_t)p + 1], s[2] * d.gp[3 * (size_t)p + 2]};
    R4<Real> q;
    q.x = (Real)(Ai[0] * g[0] + Ai[1] * g[1] + Ai[2] * g[2]);
    q.y = (Real)(Ai[3] * g[0] + Ai[4] * g[1] + Ai[5] * g[2]);
    q.z = (Real)(Ai[6] * g[0] + Ai[7] * g[1] + Ai[8] * g[2]);
    q.w = (Real)0;
    d.qp[p] = q;
}

template <typename Real>
__device__ __forceinline__ void sym3_mul(const Real* S6, const Real v[3], Real out[3]) {
    out[0] = S6[0] * v[0] + S6[1] * v[1] + S6[2] * v[2];
    out[1] = S6[1] * v[0] + S6[3] * v[1] + S6[4] * v[2];
    out[2] = S6[2] * v[0] + S6[4] * v[1] + S6[5] * v[2];
}

// Corrected Jacobians scaled by the Jacobi scaling: Jc <- Jc S_c, Jp <- Jp S_p.
template <typename Real>
__device__ __forceinline__ void scale_jacobians(const Dev<Real>& d, const Real sc[9], int p, Real Jc[18], Real Jp[6]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Jc[i] *= sc[i];
        Jc[9 + i] *= sc[i];
    }
    const Real s0 = (Real)d.sp[3 * (size_t)p], s1 = (Real)d.sp[3 * (size_t)p + 1], s2 = (Real)d.sp[3 * (size_t)p + 2];
    Jp[0] *= s0; Jp[3] *= s0;
    Jp[1] *= s1; Jp[4] *= s1;
    Jp[2] *= s2; Jp[5] *= s2;
}

// I - G for one observation, G = J~p V*^-1 J~p' (2x2), formed in FP64 with the FP64 inverse of the
// damped point block: for a weakly determined point V* is ill-conditioned and G's eigenvalues
// approach 1, where an FP32 inverse would push 1 - G negative.
template <typename Real, typename Out>
__device__ __forceinline__ void point_complement(const Dev<Real>& d, int p, const Real Jp[6], Out& A00, Out& A01,
                                                 Out& A11) {
    const double* Vi = d.Vinv64 + 6 * (size_t)p;
    const double j0[3] = {(double)Jp[0], (double)Jp[1], (double)Jp[2]};
    const double j1[3] = {(double)Jp[3], (double)Jp[4], (double)Jp[5]};
    double h0[3], h1[3];
    sym3_mul<double>(Vi, j0, h0);
    sym3_mul<double>(Vi, j1, h1);
    A00 = (Out)(1.0 - (j0[0] * h0[0] + j0[1] * h0[1] + j0[2] * h0[2]));
    A01 = (Out)(-(j0[0] * h1[0] + j0[1] * h1[1] + j0[2] * h1[2]));
    A11 = (Out)(1.0 - (j1[0] * h1[0] + j1[1] * h1[1] + j1[2] * h1[2]));
}

// One WG per camera: b_c, M_c (residual form above) and D_c^2.
template <typename Real, int M>
__global__ void __launch_bounds__(kWG) k_rhs_cam(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done) return;
    __shared__ double lds[54 * kWaves];
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    const CameraPrep<Real> P = d.prep[c];
    Real sc[9];
    for (int i = 0; i < 9; ++i) sc[i] = (Real)d.sc[9 * (size_t)c + i];
    double acc[54];
#pragma unroll
    for (int i = 0; i < 54; ++i) acc[i] = 0.0;
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6];
        corrected_jacobian<Real, M>(d, P, j, p, r, Jc, Jp);
        scale_jacobians<Real>(d, sc, p, Jc, Jp);
        const R4<Real> q = d.qp[p];
        const Real e0 = r[0] - (Jp[0] * q.x + Jp[1] * q.y + Jp[2] * q.z);
        const Real e1 = r[1] - (Jp[3] * q.x + Jp[4] * q.y + Jp[5] * q.z);
        // the camera block's per-observation product J~c'(I - G)J~c in FP64 (exactly PSD inputs)
        double A00, A01, A11;
        point_complement<Real, double>(d, p, Jp, A00, A01, A11);
        double J0[9], J1[9], H0[9], H1[9];  // (I - G) Jc~ (2x9)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            J0[k] = Jc[k];
            J1[k] = Jc[9 + k];
            H0[k] = A00 * J0[k] + A01 * J1[k];
            H1[k] = A01 * J0[k] + A11 * J1[k];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int k = i; k < 9; ++k) acc[sym9(i, k)] += J0[i] * H0[k] + J1[i] * H1[k];
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[45 + i] += (double)(Jc[i] * e0 + Jc[9 + i] * e1);
    }
    block_sum<54>(acc, lds);
    const int t = threadIdx.x;
    if (t < 81) {
        const int i = t / 9, k = t % 9;
        double m = lds[sym9(i < k ? i : k, i < k ? k : i) * kWaves];
        if (i == k) {
            const double si = d.sc[9 * (size_t)c + i];
            const double dd = fmin(fmax(si * d.U[9 * (size_t)c + i] * si, d.o.min_diag), d.o.max_diag) / st->radius;
            d.dc2[9 * (size_t)c + i] = dd;
            m += dd;
        }
        d.M[81 * (size_t)c + t] = m;
    }
    if (t < 9) d.bc[9 * (size_t)c + t] = lds[(45 + t) * kWaves];
}

constexpr int kCamsPerUpd = kWG / 16;  // camera-vector kernels: 16 threads per camera

// 16 threads per camera: thread 0 factors M_c and writes Minv_c = M_c^-1 (FP64); threads 0..8
// start the PCG: x = 0, r = b, z = Minv b, p = 0; per-block partials r.z and b.b.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_cam_factor(Dev<Real> d) {
    __shared__ double lds[2 * kWaves];
    __shared__ double Mi[kCamsPerUpd][81];
    LMState* st = d.st;
    if (st->done) return;
    const int lc = threadIdx.x >> 4, i = threadIdx.x & 15;
    const int c = blockIdx.x * kCamsPerUpd + lc;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->pcg_done = 0;
        st->pcg_iters = 0;
    }
    if (c < d.nc && i == 0) {
        double A[81];
        for (int k = 0; k < 81; ++k) A[k] = d.M[81 * (size_t)c + k];
        if (!chol_inplace<9>(A)) {
            st->step_invalid = 1;
            for (int k = 0; k < 81; ++k) Mi[lc][k] = 0.0;
        } else {
            double Ai[81];
            chol_inverse<9>(A, Ai);
            for (int k = 0; k < 81; ++k) Mi[lc][k] = Ai[k];
        }
    }
    __syncthreads();
    double rz = 0.0, bb = 0.0;
    if (c < d.nc && i < 9) {
        const size_t o = 9 * (size_t)c;
        double z = 0.0;
        for (int k = 0; k < 9; ++k) z += Mi[lc][9 * i + k] * d.bc[o + k];
        for (int k = 0; k < 9; ++k) d.Minv[81 * (size_t)c + 9 * i + k] = Mi[lc][9 * i + k];
        const double b = d.bc[o + i];
        d.pz[o + i] = z;
        d.pr[o + i] = b;
        d.px[o + i] = 0.0;
        d.pp[o + i] = 0.0;
        rz = b * z;
        bb = b * b;
    }
    double a[2] = {rz, bb};
    block_sum<2>(a, lds);
    if (threadIdx.x == 0) {
        d.upd_part[blockIdx.x] = lds[0];
        d.upd_part[d.n_upd + blockIdx.x] = lds[kWaves];
    }
}

// ---------------------------------------------------------------------------------------------
// Implicit-Schur PCG, four kernels per iteration it = 1, 2, ...:
//   k_jtx(it)        beta and the convergence test from the previous update's partials (every
//                    WG reduces the same few partials in the same order, so all agree), then
//                    p <- z + beta p and u_j = Jp'(Jc S_c p_c) per observation;
//   k_pt_z           z_p = V*^-1 S_p sum u_j, zs_p = S_p z_p, partial z'D_p^2 z;
//   k_jz             q_c = D_c^2 p_c + S_c sum Jc' e_j, partial sum |e|^2 + p'D_c^2 p;
//   k_cam_update(it) alpha = r.z / p'Sp, x += alpha p, r -= alpha q, z = Minv r, partials.
// Stop: ||r|| <= tol ||b|| or pcg_max_iterations (or a non-positive p'Sp, i.e. rounding level).

template <int NW = kWaves>
__device__ __forceinline__ void reduce_upd(const double* __restrict__ part, int n, double& a, double& b, double* lds) {
    double v[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += 64 * NW) {
        v[0] += part[i];
        v[1] += part[n + i];
    }
    block_sum<2, double, NW>(v, lds);
    a = lds[0];
    b = lds[NW];
}

// Start of PCG iteration `it` (>= 1) in every workgroup: the convergence test and beta from the
// previous update's partials (each workgroup reduces them in the same order, so all agree).
// Returns true when the iteration does not run.
template <typename Real, int NW = kWaves>
__device__ __forceinline__ bool pcg_beta(const Dev<Real>& d, int it, double* lds, double& beta) {
    LMState* st = d.st;
    double rz, rr;
    reduce_upd<NW>(d.upd_part, d.n_upd, rz, rr, lds);
    bool done;
    if (it == 1)
        done = !(rr > 0.0) || d.o.pcg_max_iterations <= 0;
    else
        done = rr <= d.o.pcg_tol * d.o.pcg_tol * st->pcg_bb || it - 1 >= d.o.pcg_max_iterations;
    if (!isfinite(rz) || !isfinite(rr)) done = true;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (it == 1) st->pcg_bb = rr;
        d.rz_hist[it] = rz;
        st->pcg_iters = done ? it - 1 : it;
        if (done) st->pcg_done = 1;
        if (!isfinite(rz) || !isfinite(rr)) st->step_invalid = 1;
    }
    beta = done || it == 1 ? 0.0 : rz / d.rz_hist[it - 1];
    return done;
}

// One WG per camera.  mode 1 (PCG iteration it): see above.  mode 0 (back-substitution):
// u_j = Jp'(r~ - Jc S_c y_c) with y = vsrc.
template <typename Real, int M>
__global__ void __launch_bounds__(kWG) k_jtx(Dev<Real> d, int mode, int it, const double* __restrict__ vsrc) {
    LMState* st = d.st;
    if (st->done || st->step_invalid || (mode == 1 && st->pcg_done)) return;
    __shared__ double lds[2 * kWaves];
    __shared__ double vsh[9];
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    // the first observation's inputs are on their way while beta is reduced
    const int j_end = d.cam_off[c + 1];
    int j = d.cam_off[c] + threadIdx.x;
    ObsIn<Real> in{};
    if (j < j_end) load_obs<Real, M>(d, j, in);
    // the camera's prep and scaling are requested before the reduction and the barrier
    const CameraPrep<Real> P = d.prep[c];
    double scl[9];
    for (int i = 0; i < 9; ++i) scl[i] = d.sc[9 * (size_t)c + i];
    double beta = 0.0;
    if (mode == 1 && pcg_beta(d, it, lds, beta)) return;
    if (threadIdx.x < 9) {
        const size_t i = 9 * (size_t)c + threadIdx.x;
        double v = vsrc[i];
        if (mode == 1) {
            v = d.pz[i] + beta * v;
            d.pp[i] = v;
        }
        vsh[threadIdx.x] = v;
    }
    __syncthreads();
    Real vs[9];
    for (int i = 0; i < 9; ++i) vs[i] = (Real)(scl[i] * vsh[i]);
    for (; j < j_end; j += kWG) {
        const ObsIn<Real> cur = in;
        if (j + kWG < j_end) load_obs<Real, M>(d, j + kWG, in);
        Real r[2], Jc[18], Jp[6];
        corrected_from<Real, M>(d, P, cur, r, Jc, Jp);
        Real a0 = 0, a1 = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            a0 += Jc[i] * vs[i];
            a1 += Jc[9 + i] * vs[i];
        }
        if (mode == 0) {
            a0 = r[0] - a0;
            a1 = r[1] - a1;
        }
        R4<Real> u;
        u.x = Jp[0] * a0 + Jp[3] * a1;
        u.y = Jp[1] * a0 + Jp[4] * a1;
        u.z = Jp[2] * a0 + Jp[5] * a1;
        u.w = (Real)0;
        d.u[d.slot_of[j]] = u;
    }
}

// One workgroup per chunk: t_p = S_p sum u_j, z_p = Vinv t_p, zs_p = S_p z_p; partial z'D_p^2 z.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_pt_z(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid || st->pcg_done) return;
    __shared__ double lds[(kWaves - 1) * 3 * 64];
    const int ch = xcd_remap(blockIdx.x, d.n_chunks);
    if (ch >= d.n_chunks) return;
    // the point's scaling, V*^-1 and D_p^2 are requested before the sum (wave 0 uses them)
    const int p = 64 * ch + threadIdx.x;
    const bool own = threadIdx.x < 64 && p < d.np;
    double s[3], Vi[6], dp[3];
    for (int i = 0; i < 3; ++i) {
        s[i] = own ? d.sp[3 * (size_t)p + i] : 0.0;
        dp[i] = own ? d.dp2[3 * (size_t)p + i] : 0.0;
    }
    for (int i = 0; i < 6; ++i) Vi[i] = own ? d.Vinv64[6 * (size_t)p + i] : 0.0;
    double a[3];
    chunk_sum_r4<Real>(d.chunk_len[ch], (size_t)d.chunk_off[ch], d.u, a, lds);
    if (threadIdx.x >= 64) return;
    double zdz = 0.0;
    if (own) {
        double t[3], z[3];
        for (int i = 0; i < 3; ++i) t[i] = s[i] * a[i];
        sym3_mul<double>(Vi, t, z);  // the point block solve stays FP64
        d.zs[p] = make_double4(s[0] * z[0], s[1] * z[1], s[2] * z[2], 0.0);
        for (int i = 0; i < 3; ++i) zdz += z[i] * dp[i] * z[i];
    }
    zdz = wave_sum(zdz);
    if (threadIdx.x == 0) d.pt_part[PS_B * (size_t)d.n_pt_blocks + ch] = zdz;
}

// One WG per camera: q_c = D_c^2 p_c + S_c sum Jc' e_j, e_j = Jc (S_c p_c) - Jp zs_p;
// partial sum |e|^2 + p_c'D_c^2 p_c.
template <typename Real, int M>
__global__ void __launch_bounds__(kWG) k_jz(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid || st->pcg_done) return;
    __shared__ double lds[10 * kWaves];
    __shared__ double psh[9];
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    const int j_end = d.cam_off[c + 1];
    int j = d.cam_off[c] + threadIdx.x;
    ObsIn<Real> in{};
    if (j < j_end) load_obs<Real, M>(d, j, in);
    // the camera's prep and scaling are requested before the barrier
    const CameraPrep<Real> P = d.prep[c];
    double scl[9];
    for (int i = 0; i < 9; ++i) scl[i] = d.sc[9 * (size_t)c + i];
    if (threadIdx.x < 9) psh[threadIdx.x] = d.pp[9 * (size_t)c + threadIdx.x];
    __syncthreads();
    Real ps[9];
    for (int i = 0; i < 9; ++i) ps[i] = (Real)(scl[i] * psh[i]);
    double acc[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) acc[i] = 0.0;
    for (; j < j_end; j += kWG) {
        const ObsIn<Real> cur = in;
        if (j + kWG < j_end) load_obs<Real, M>(d, j + kWG, in);
        Real r[2], Jc[18], Jp[6];
        corrected_from<Real, M>(d, P, cur, r, Jc, Jp);
        const double4 z4 = d.zs[cur.p];
        const Real z[3] = {(Real)z4.x, (Real)z4.y, (Real)z4.z};
        Real e0 = -(Jp[0] * z[0] + Jp[1] * z[1] + Jp[2] * z[2]);
        Real e1 = -(Jp[3] * z[0] + Jp[4] * z[1] + Jp[5] * z[2]);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            e0 += Jc[i] * ps[i];
            e1 += Jc[9 + i] * ps[i];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += (double)(Jc[i] * e0 + Jc[9 + i] * e1);
        acc[9] += (double)e0 * (double)e0 + (double)e1 * (double)e1;
    }
    block_sum<10>(acc, lds);
    const int t = threadIdx.x;
    if (t < 9) {
        const size_t i = 9 * (size_t)c + t;
        d.qc[i] = d.dc2[i] * psh[t] + d.sc[i] * lds[t * kWaves];
    }
    if (t == 0) {
        double s = lds[9 * kWaves];
        for (int i = 0; i < 9; ++i) s += psh[i] * d.dc2[9 * (size_t)c + i] * psh[i];
        d.cam_part[PS_A * (size_t)d.nc + c] = s;
    }
}

// The two camera-order PCG products of the Snavely model (Ceres additive update) in direction form
// (snavely.h, rot_directional): the same u_j and q_c as k_jtx (mode 1) and k_jz, without forming
// the 2x9 / 2x3 Jacobian per observation.  The Huber weight sqrt(rho') of each observation and its
// point relative to the camera centre were stored by k_lin_cam at the linearization point
// (co_rec), so neither the measurement nor the point is re-read.

// Per-camera constants of a direction-form product: dR_v for the scaled direction v = S_c p_c,
// the rotation matrix and v's translation / intrinsic parts in working precision.  The rotation
// columns are taken about the camera centre c = -R't,
//   dR_v X + v_t = dR_v (X - c) + (dR_v c + v_t),
// with X - c (the observation record, formed in FP64 at the linearization) and dR_v c + v_t formed
// in FP64: for a point near a camera far from the world origin, dR_v X and v_t are large and nearly
// opposite, and in FP32 their sum would keep few digits.
template <typename Real>
struct DirCam {
    Real Rd[9], R[9], vt[3], vf, vk1, vk2;
};

// The jz product's difference vector dR_v (X - c) + (dR_v c + v_t) - R zs: the camera's motion and
// the point's, seen from the camera, nearly cancel where the point follows the camera, so in mixed
// precision it is formed in FP64 (from the record's X - c and the FP64 zs) and rounded once.
template <typename Real>
__device__ __forceinline__ void jz_difference(const CameraPrep<Real>& P, const double* rsh, const double vt64[3],
                                              const R4<Real>& rec, const double4& z, Real dq[3]) {
    const double X[3] = {(double)rec.x, (double)rec.y, (double)rec.z};
    const double Z[3] = {z.x, z.y, z.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double q = rsh[3 * i] * X[0] + rsh[3 * i + 1] * X[1] + rsh[3 * i + 2] * X[2] + vt64[i];
        const double zr = P.R64[3 * i] * Z[0] + P.R64[3 * i + 1] * Z[1] + P.R64[3 * i + 2] * Z[2];
        dq[i] = (Real)(q - zr);
    }
}

// A value that is the same in every lane, moved to a scalar register (the per-camera constants of
// the direction-form kernels would otherwise occupy ~24 vector registers of every lane).
__device__ __forceinline__ float wave_uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double wave_uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename Real>
__device__ __forceinline__ void load_dircam(const CameraPrep<Real>& P, const double* vsh, const double* rsh,
                                            DirCam<Real>& C) {
    double c[3];
    camera_centre<Real>(P, c);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        C.Rd[k] = wave_uniform((Real)rsh[k]);
        C.R[k] = wave_uniform((Real)P.R64[k]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        C.vt[i] = wave_uniform((Real)(vsh[3 + i] + rsh[3 * i] * c[0] + rsh[3 * i + 1] * c[1] + rsh[3 * i + 2] * c[2]));
    C.vf = wave_uniform((Real)vsh[6]);
    C.vk1 = wave_uniform((Real)vsh[7]);
    C.vk2 = wave_uniform((Real)vsh[8]);
}

// The projection's linearization from the record's X - c: p = R (X - c) in working precision.
template <typename Real>
__device__ __forceinline__ void proj_lin_rec(const CameraPrep<Real>& P, const DirCam<Real>& C, const R4<Real>& rec,
                                             ProjLin<Real>& L) {
    Real p[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) p[i] = C.R[3 * i] * rec.x + C.R[3 * i + 1] * rec.y + C.R[3 * i + 2] * rec.z;
    L.ginv = (Real)1 / p[2];
    L.xp = -p[0] * L.ginv;
    L.yp = -p[1] * L.ginv;
    L.r2 = L.xp * L.xp + L.yp * L.yp;
    const Real poly = P.k1 + P.k2 * L.r2;
    L.dist = (Real)1 + L.r2 * poly;
    L.fd = P.f * L.dist;
    L.fe = (Real)2 * P.f * (poly + P.k2 * L.r2);
}

// a = Jc v / w - D R zR... for one observation record (w: its Huber weight): D (dR_v (X - c) +
// dR_v c + v_t - zR) + K (v_f, v_k1, v_k2), and the projection's linearization.
template <typename Real>
__device__ __forceinline__ void dir_apply(const CameraPrep<Real>& P, const DirCam<Real>& C, const R4<Real>& rec,
                                          const Real zR[3], ProjLin<Real>& L, Real a[2]) {
    proj_lin_rec<Real>(P, C, rec, L);
    Real q[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        q[i] = C.Rd[3 * i] * rec.x + C.Rd[3 * i + 1] * rec.y + C.Rd[3 * i + 2] * rec.z + C.vt[i] - zR[i];
    proj_apply<Real>(L, q, a);
    const Real kap = L.dist * C.vf + P.f * L.r2 * (C.vk1 + L.r2 * C.vk2);
    a[0] += L.xp * kap;
    a[1] += L.yp * kap;
}

// D dq + K (v_f, v_k1, v_k2) for a given camera-frame vector dq, and the projection's linearization.
template <typename Real>
__device__ __forceinline__ void dir_apply_dq(const CameraPrep<Real>& P, const DirCam<Real>& C, const R4<Real>& rec,
                                             const Real dq[3], ProjLin<Real>& L, Real a[2]) {
    proj_lin_rec<Real>(P, C, rec, L);
    proj_apply<Real>(L, dq, a);
    const Real kap = L.dist * C.vf + P.f * L.r2 * (C.vk1 + L.r2 * C.vk2);
    a[0] += L.xp * kap;
    a[1] += L.yp * kap;
}

// PCG iteration `it`: beta, p <- z + beta p, then u_j = Jp'(Jc S_c p_c) per observation (slot order).
// An observation needs only its record and slot (no dependent loads).
template <typename Real, int NT = kWG>
__global__ void __launch_bounds__(NT) k_jtx_dir(Dev<Real> d, int it) {
    constexpr int NW = NT / 64;
    LMState* st = d.st;
    if (st->done || st->step_invalid || st->pcg_done) return;
    __shared__ double lds[2 * NW];
    __shared__ double vsh[9], rsh[9];
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    const int j_end = d.cam_off[c + 1];
    int j = d.cam_off[c] + threadIdx.x;
    R4<Real> rec{};
    int slot = 0;
    if (j < j_end) {
        rec = d.co_rec[j];
        slot = d.slot_of[j];
    }
    double pz = 0.0, pp = 0.0, sc = 0.0;
    if (threadIdx.x < 9) {
        const size_t i = 9 * (size_t)c + threadIdx.x;
        pz = d.pz[i];
        pp = d.pp[i];
        sc = d.sc[i];
    }
    // the camera's prep is requested before the reductions (a load after a barrier waits for it)
    const CameraPrep<Real> P = d.prep[c];
    double beta;
    if (pcg_beta<Real, NW>(d, it, lds, beta)) return;
    if (threadIdx.x < 9) {
        const double v = pz + beta * pp;
        d.pp[9 * (size_t)c + threadIdx.x] = v;
        vsh[threadIdx.x] = sc * v;
    }
    __syncthreads();
    if (threadIdx.x == 0) rot_directional<Real>(P, vsh, rsh);
    __syncthreads();
    DirCam<Real> C;
    load_dircam<Real>(P, vsh, rsh, C);
    const Real zero[3] = {(Real)0, (Real)0, (Real)0};
    for (; j < j_end; j += NT) {
        const R4<Real> cur = rec;
        const int cur_slot = slot;
        if (j + NT < j_end) {
            rec = d.co_rec[j + NT];
            slot = d.slot_of[j + NT];
        }
        ProjLin<Real> L;
        Real a[2];
        dir_apply<Real>(P, C, cur, zero, L, a);
        const Real ww = cur.w * cur.w;
        a[0] *= ww;
        a[1] *= ww;
        Real g[3];
        proj_apply_t<Real>(L, a, g);
        R4<Real> u;  // R' g
        u.x = C.R[0] * g[0] + C.R[3] * g[1] + C.R[6] * g[2];
        u.y = C.R[1] * g[0] + C.R[4] * g[1] + C.R[7] * g[2];
        u.z = C.R[2] * g[0] + C.R[5] * g[1] + C.R[8] * g[2];
        u.w = (Real)0;
        d.u[cur_slot] = u;
    }
}

// q_c = D_c^2 p_c + S_c sum Jc' e_j with e_j = Jc S_c p_c - Jp zs_p; partial sum |e|^2 + p'D_c^2 p.
// Per observation only g = D'(w e) is formed (one gather: the point's zs); the camera accumulates
// G = sum g (X - c)', sum g and the intrinsic sums, and contracts G + (sum g) c' with dR/dw once.
template <typename Real, int NT = kWG>
__global__ void __launch_bounds__(NT) k_jz_dir(Dev<Real> d) {
    constexpr int NW = NT / 64;
    LMState* st = d.st;
    if (st->done || st->step_invalid || st->pcg_done) return;
    __shared__ double lds[16 * NW];
    __shared__ double psh[9], vsh[9], rsh[9], qsh[9], dsh[9];
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    const int j_end = d.cam_off[c + 1];
    int j = d.cam_off[c] + threadIdx.x;
    // loads in flight ahead of the products: the point index two rounds ahead, the record and the
    // point's z (a gather through that index) one round ahead
    R4<Real> rec{};
    double4 zin{};
    int pt_next = 0;
    if (j < j_end) {
        rec = d.co_rec[j];
        zin = d.zs[d.co_pt[j]];
    }
    if (j + NT < j_end) pt_next = d.co_pt[j + NT];
    const CameraPrep<Real> P = d.prep[c];  // requested before the barrier (a load after it waits for it)
    double sc_t = 0.0, dc2_t = 0.0;          // thread t < 9: S_c and D_c^2 entry t, kept for the end
    if (threadIdx.x < 9) {
        const size_t i = 9 * (size_t)c + threadIdx.x;
        sc_t = d.sc[i];
        dc2_t = d.dc2[i];
        psh[threadIdx.x] = d.pp[i];
        vsh[threadIdx.x] = sc_t * d.pp[i];
        dsh[threadIdx.x] = dc2_t;
    }
    __syncthreads();
    if (threadIdx.x == 0) rot_directional<Real>(P, vsh, rsh);
    __syncthreads();
    DirCam<Real> C;
    load_dircam<Real>(P, vsh, rsh, C);
    double vt64[3];  // dR_v c + v_t in FP64
    {
        double cc[3];
        camera_centre<Real>(P, cc);
        for (int i = 0; i < 3; ++i)
            vt64[i] = wave_uniform(vsh[3 + i] + rsh[3 * i] * cc[0] + rsh[3 * i + 1] * cc[1] + rsh[3 * i + 2] * cc[2]);
    }
    double acc[16];  // G (row-major 3x3), sum g, intrinsic sums (dist, r2, r2^2 times x'e0 + y'e1), |e|^2
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0;
    for (; j < j_end; j += NT) {
        const R4<Real> cur = rec;
        const double4 z = zin;
        if (j + NT < j_end) {
            rec = d.co_rec[j + NT];
            zin = d.zs[pt_next];
            if (j + 2 * NT < j_end) pt_next = d.co_pt[j + 2 * NT];
        }
        Real dq[3];
        jz_difference<Real>(P, rsh, vt64, cur, z, dq);
        ProjLin<Real> L;
        Real a[2];
        dir_apply_dq<Real>(P, C, cur, dq, L, a);
        const Real e0 = cur.w * a[0], e1 = cur.w * a[1];
        acc[15] += (double)e0 * (double)e0 + (double)e1 * (double)e1;
        const Real ew[2] = {cur.w * e0, cur.w * e1};
        Real g[3];
        proj_apply_t<Real>(L, ew, g);
        const double X[3] = {(double)cur.x, (double)cur.y, (double)cur.z};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double gi = (double)g[i];
#pragma unroll
            for (int m = 0; m < 3; ++m) acc[3 * i + m] += gi * X[m];
            acc[9 + i] += gi;
        }
        const double se = (double)(L.xp * ew[0] + L.yp * ew[1]);
        const double r2 = (double)L.r2;
        acc[12] += (double)L.dist * se;
        acc[13] += r2 * se;
        acc[14] += r2 * r2 * se;
    }
    block_sum<16, double, NW>(acc, lds);
    const int t = threadIdx.x;
    if (t == 0) {
        double cc[3], G[9], w3[3];
        camera_centre<Real>(P, cc);
        const double gs[3] = {lds[9 * NW], lds[10 * NW], lds[11 * NW]};
        for (int i = 0; i < 3; ++i)
            for (int m = 0; m < 3; ++m) G[3 * i + m] = lds[(3 * i + m) * NW] + gs[i] * cc[m];
        rot_contract<Real>(P, G, w3);
        const double f = (double)P.f;
        const double qv[9] = {w3[0], w3[1], w3[2], gs[0], gs[1], gs[2],
                              lds[12 * NW], f * lds[13 * NW], f * lds[14 * NW]};
        for (int k = 0; k < 9; ++k) qsh[k] = qv[k];
    }
    __syncthreads();
    if (t < 9) d.qc[9 * (size_t)c + t] = dc2_t * psh[t] + sc_t * qsh[t];
    if (t == 0) {
        double s = lds[15 * NW];
        for (int i = 0; i < 9; ++i) s += psh[i] * dsh[i] * psh[i];
        d.cam_part[PS_A * (size_t)d.nc + c] = s;
    }
}

// 16 threads per camera: alpha = rz / p'Sp (each block reduces the same partials in the same
// order), x += alpha p, r -= alpha q, z = Minv r (coalesced Minv rows); per-block partials r.z
// and r.r for k_jtx.  The camera's vectors and Minv row do not depend on alpha: their loads are
// issued before the reduction, so the kernel waits on memory once.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_cam_update(Dev<Real> d, int it) {
    LMState* st = d.st;
    if (st->done || st->step_invalid || st->pcg_done) return;
    __shared__ double lds[2 * kWaves];
    __shared__ double rsh[kCamsPerUpd][9];
    const int lc = threadIdx.x >> 4, i = threadIdx.x & 15;
    const int c = blockIdx.x * kCamsPerUpd + lc;
    const bool act = c < d.nc && i < 9;
    const size_t k = act ? 9 * (size_t)c + i : 0;
    double x = 0.0, p = 0.0, r = 0.0, q = 0.0, Mrow[9];
    if (act) {
        x = d.px[k];
        p = d.pp[k];
        r = d.pr[k];
        q = d.qc[k];
    }
    const double* Mi = d.Minv + (act ? 81 * (size_t)c + 9 * i : 0);
#pragma unroll
    for (int m = 0; m < 9; ++m) Mrow[m] = act ? Mi[m] : 0.0;
    const double rz_it = d.rz_hist[it];
    double v = strided_sum(d.cam_part + PS_A * (size_t)d.nc, d.nc);
    v += strided_sum(d.pt_part + PS_B * (size_t)d.n_pt_blocks, d.n_pt_blocks);
    const double pSp = block_sum1(v, lds);
    if (!(pSp > 0.0) || !isfinite(pSp)) {
        // v'Sv is a sum of squares: zero means the direction vanished (rounding level)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->pcg_done = 1;
            st->pcg_iters = it - 1;
        }
        return;
    }
    const double alpha = rz_it / pSp;
    if (act) {
        d.px[k] = x + alpha * p;
        r = r - alpha * q;
        d.pr[k] = r;
        rsh[lc][i] = r;
    }
    __syncthreads();
    double rz = 0.0, rr = 0.0;
    if (act) {
        double z = 0.0;
        for (int m = 0; m < 9; ++m) z += Mrow[m] * rsh[lc][m];
        d.pz[k] = z;
        rz = r * z;
        rr = r * r;
    }
    double a[2] = {rz, rr};
    block_sum<2>(a, lds);
    if (threadIdx.x == 0) {
        d.upd_part[blockIdx.x] = lds[0];
        d.upd_part[d.n_upd + blockIdx.x] = lds[kWaves];
    }
}

// ---------------------------------------------------------------------------------------------
// Step, candidate and decision

// One workgroup per chunk: y_p = Vinv S_p sum u_j (u_j = Jp'(r~ - Jc S_c y_c)),
// delta_p = -S_p y_p, candidate = x + delta.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_pt_backsub(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid) return;
    __shared__ double lds[(kWaves - 1) * 3 * 64];
    const int ch = xcd_remap(blockIdx.x, d.n_chunks);
    if (ch >= d.n_chunks) return;
    double a[3];
    chunk_sum<3>(d.chunk_len[ch], (size_t)d.chunk_off[ch], a,
                 [&](size_t slot, double (&acc)[3]) {
                     const R4<Real> u = d.u[slot];
                     acc[0] += (double)u.x;
                     acc[1] += (double)u.y;
                     acc[2] += (double)u.z;
                 },
                 lds);
    if (threadIdx.x >= 64) return;
    const int p = 64 * ch + threadIdx.x;
    double dx2 = 0.0, x2 = 0.0;
    if (p < d.np) {
        double s[3], t[3];
        for (int i = 0; i < 3; ++i) {
            s[i] = d.sp[3 * (size_t)p + i];
            t[i] = s[i] * a[i];
        }
        const double* Vi = d.Vinv64 + 6 * (size_t)p;
        const double y[3] = {Vi[0] * t[0] + Vi[1] * t[1] + Vi[2] * t[2], Vi[1] * t[0] + Vi[3] * t[1] + Vi[4] * t[2],
                             Vi[2] * t[0] + Vi[4] * t[1] + Vi[5] * t[2]};
        for (int i = 0; i < 3; ++i) {
            const double dl = -s[i] * y[i];
            const double x = d.pt_x[4 * (size_t)p + i];
            d.dpt[4 * (size_t)p + i] = dl;
            d.pt_c[4 * (size_t)p + i] = x + dl;
            dx2 += dl * dl;
            x2 += x * x;
        }
    }
    dx2 = wave_sum(dx2);
    x2 = wave_sum(x2);
    if (threadIdx.x == 0) {
        d.pt_part[PS_DX2 * (size_t)d.n_pt_blocks + ch] = dx2;
        d.pt_part[PS_X2 * (size_t)d.n_pt_blocks + ch] = x2;
    }
}

// One thread per camera: delta_c = -S_c y_c, candidate camera and its FP64 prep.
template <typename Real, int M>
__global__ void __launch_bounds__(64) k_cam_step(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.nc) return;
    double cam[9], dx2 = 0.0, x2 = 0.0;
    for (int i = 0; i < 9; ++i) {
        const double dl = -d.sc[9 * (size_t)c + i] * d.px[9 * (size_t)c + i];
        const double x = d.cam_x[(size_t)i * d.nc + c];
        d.dcam[(size_t)i * d.nc + c] = dl;
        cam[i] = x + dl;
        dx2 += dl * dl;
        x2 += x * x;
    }
    if constexpr (M == kModelStereo) {
        // pose <- exp(xi) pose, xi = (rho, phi) in the first six slots
        double xi[6], w[3], t[3];
        for (int i = 0; i < 6; ++i) xi[i] = d.dcam[(size_t)i * d.nc + c];
        for (int i = 0; i < 3; ++i) {
            w[i] = d.cam_x[(size_t)i * d.nc + c];
            t[i] = d.cam_x[(size_t)(3 + i) * d.nc + c];
        }
        se3_left_update(xi, w, t, cam, cam + 3);
    } else if constexpr (M == kModelSnavelyG2o) {
        // rotation <- exp(delta) rotation, read back as angle-axis
        double dd[3], w[3];
        for (int i = 0; i < 3; ++i) {
            dd[i] = d.dcam[(size_t)i * d.nc + c];
            w[i] = d.cam_x[(size_t)i * d.nc + c];
        }
        so3_left_update(dd, w, cam);
    }
    for (int i = 0; i < 9; ++i) d.cam_c[(size_t)i * d.nc + c] = cam[i];
    CameraPrep<double> P;
    camera_prep<double>(cam, P);
    d.prep_c[c] = P;
    d.cam_part[PS_DX2 * (size_t)d.nc + c] = dx2;
    d.cam_part[PS_X2 * (size_t)d.nc + c] = x2;
}

// One WG per camera: model cost change -(J~ s).(r~ + J~ s / 2) with J, r at x (working
// precision), and the FP64 candidate cost at x + delta.
template <typename Real, int M>
__global__ void __launch_bounds__(kWG) k_model_cand(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid) return;
    __shared__ double lds[2 * kWaves];
    const int c = xcd_remap(blockIdx.x, d.nc);
    if (c >= d.nc) return;
    const CameraPrep<Real> P = d.prep[c];
    const CameraPrep<double> Pc = d.prep_c[c];
    Real dc[9];
    for (int i = 0; i < 9; ++i) dc[i] = (Real)d.dcam[(size_t)i * d.nc + c];
    double acc[2] = {0.0, 0.0};
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6];
        corrected_jacobian<Real, M>(d, P, j, p, r, Jc, Jp);
        const double4 dp4 = reinterpret_cast<const double4*>(d.dpt)[p];
        const Real dp[3] = {(Real)dp4.x, (Real)dp4.y, (Real)dp4.z};
        Real m0 = Jp[0] * dp[0] + Jp[1] * dp[1] + Jp[2] * dp[2];
        Real m1 = Jp[3] * dp[0] + Jp[4] * dp[1] + Jp[5] * dp[2];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            m0 += Jc[i] * dc[i];
            m1 += Jc[9 + i] * dc[i];
        }
        acc[0] += (double)m0 * ((double)r[0] + 0.5 * (double)m0) + (double)m1 * ((double)r[1] + 0.5 * (double)m1);
        double X[3];
        load_point<double>(d.pt_c, p, X);
        double rc[2];
        if constexpr (M == kModelStereo)
            stereo_residual<double>(Pc, X, d.rig, d.co_side[j], d.co_uv64[2 * (size_t)j], d.co_uv64[2 * (size_t)j + 1], rc);
        else
            snavely_residual<double>(Pc, X, d.co_uv64[2 * (size_t)j], d.co_uv64[2 * (size_t)j + 1], rc);
        acc[1] += half_rho(rc[0] * rc[0] + rc[1] * rc[1], d.o.loss, d.o.loss_a);
    }
    block_sum<2>(acc, lds);
    if (threadIdx.x == 0) {
        d.cam_part[PS_MODEL * (size_t)d.nc + c] = -lds[0];
        d.cam_part[PS_COST * (size_t)d.nc + c] = lds[kWaves];
    }
}

// Single WG: reduce the step partials and run one Ceres-style trust-region decision.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_decide(Dev<Real> d) {
    __shared__ double lds[4 * kWaves];
    LMState* st = d.st;
    if (st->done) return;
    double a[4] = {0.0, 0.0, 0.0, 0.0};  // model, cost, |dx|^2, |x|^2
    const bool have_step = !st->step_invalid;
    if (have_step) {
        for (int i = threadIdx.x; i < d.nc; i += kWG) {
            a[0] += d.cam_part[PS_MODEL * (size_t)d.nc + i];
            a[1] += d.cam_part[PS_COST * (size_t)d.nc + i];
            a[2] += d.cam_part[PS_DX2 * (size_t)d.nc + i];
            a[3] += d.cam_part[PS_X2 * (size_t)d.nc + i];
        }
        for (int i = threadIdx.x; i < d.n_pt_blocks; i += kWG) {
            a[2] += d.pt_part[PS_DX2 * (size_t)d.n_pt_blocks + i];
            a[3] += d.pt_part[PS_X2 * (size_t)d.n_pt_blocks + i];
        }
    }
    block_sum<4>(a, lds);
    if (threadIdx.x != 0) return;
    const Opts& o = d.o;
    const double model = lds[0];
    double cand = lds[kWaves];
    const double dx2 = lds[2 * kWaves], x2 = lds[3 * kWaves];
    st->iteration++;
    ba_iteration row{};
    row.iteration = st->iteration;
    row.linear_solver_iterations = st->pcg_iters;
    st->total_pcg_iters += st->pcg_iters;
    row.model_cost_change = model;
    st->accepted = 0;
    if (o.g2o) {
        // g2o OptimizationAlgorithmLevenberg: rho = (chi2 - chi2_new) / (x'(lambda x + b) + 1e-3) with
        // chi2 = 2 cost and x'(lambda x + b) = 2 model for the solved damped system; accept when
        // rho > 0, lambda *= max(1/3, min(2/3, 1 - (2 rho - 1)^3)), else lambda *= nu, nu *= 2; an
        // iteration ends with an accepted trial, the solve after g2o_max_trials rejected ones
        // (radius = 1 / lambda, decrease_factor = nu)
        if (!isfinite(cand)) cand = DBL_MAX;
        const double step_norm = sqrt(dx2);
        row.step_norm = have_step ? step_norm : 0.0;
        row.step_is_valid = have_step ? 1 : 0;
        const double chi_change = have_step ? 2.0 * (st->x_cost - cand) : -DBL_MAX;
        const double rho = have_step ? chi_change / (2.0 * model + 1e-3) : -1.0;
        row.cost_change = have_step ? st->x_cost - cand : 0.0;
        row.relative_decrease = rho;
        st->step_invalid = 0;
        if (have_step && rho > 0.0 && cand < DBL_MAX) {
            const double t = 2.0 * rho - 1.0;
            st->radius = st->radius / fmax(1.0 / 3.0, fmin(2.0 / 3.0, 1.0 - t * t * t));
            st->decrease_factor = 2.0;
            st->x_cost = cand;
            st->accepted = 1;
            st->need_lin = 1;
            st->g2o_iter++;
            st->g2o_trials = 0;
            row.step_is_successful = 1;
            row.cost = cand;
            st->num_successful++;
        } else {
            st->radius = st->radius / st->decrease_factor;
            st->decrease_factor *= 2.0;
            st->g2o_trials++;
            row.step_is_successful = 0;
            row.cost = have_step ? cand : st->x_cost;
            st->num_unsuccessful++;
        }
        row.gradient_max_norm = st->gmax;
        row.trust_region_radius = st->radius;
        row.cumulative_time_in_seconds = lm_ticks(st);
        if (st->trace_len < d.trace_cap) d.trace[st->trace_len] = row;
        st->trace_len++;
        if (st->g2o_iter >= o.max_iterations)
            lm_finish(st, BA_NO_CONVERGENCE, BA_REASON_MAX_ITERATIONS, st->g2o_iter);
        else if (st->g2o_trials >= o.g2o_max_trials || rho == 0.0 || !isfinite(st->radius) || !(st->radius > 0.0))
            lm_finish(st, BA_CONVERGENCE, BA_REASON_G2O_STOP, rho);
        return;
    }
    const bool valid = have_step && model > 0.0 && isfinite(dx2);
    if (!valid) {
        st->step_invalid = 0;
        if (++st->consecutive_invalid >= o.max_invalid) {
            lm_finish(st, BA_FAILURE, BA_REASON_INVALID_STEPS, st->consecutive_invalid);
            return;
        }
        st->radius = st->radius / st->decrease_factor;
        st->decrease_factor *= 2.0;
        row.step_is_valid = 0;
        row.step_is_successful = 0;
        row.cost = st->x_cost;
        row.gradient_max_norm = st->gmax;
        row.trust_region_radius = st->radius;
        st->num_unsuccessful++;
    } else {
        st->consecutive_invalid = 0;
        row.step_is_valid = 1;
        if (!isfinite(cand)) cand = DBL_MAX;
        const double step_norm = sqrt(dx2), x_norm = sqrt(x2);
        row.step_norm = step_norm;
        if (step_norm <= o.ptol * (x_norm + o.ptol)) {
            lm_finish(st, BA_CONVERGENCE, BA_REASON_PARAMETER_TOLERANCE, step_norm / (x_norm + o.ptol));
            return;
        }
        const double cost_change = st->x_cost - cand;
        row.cost_change = cost_change;
        if (fabs(cost_change) <= o.ftol * st->x_cost) {
            lm_finish(st, BA_CONVERGENCE, BA_REASON_FUNCTION_TOLERANCE, fabs(cost_change) / st->x_cost);
            return;
        }
        const double rho = cost_change / model;
        row.relative_decrease = rho;
        if (rho > o.min_rel_decrease) {
            st->x_cost = cand;
            const double t = 2.0 * rho - 1.0;
            st->radius = fmin(o.max_radius, st->radius / fmax(1.0 / 3.0, 1.0 - t * t * t));
            st->decrease_factor = 2.0;
            st->accepted = 1;
            st->need_lin = 1;
            row.step_is_successful = 1;
            row.cost = cand;
            st->num_successful++;
        } else {
            st->radius = st->radius / st->decrease_factor;
            st->decrease_factor *= 2.0;
            row.step_is_successful = 0;
            row.cost = cand;
            st->num_unsuccessful++;
        }
        row.gradient_max_norm = st->gmax;
        row.trust_region_radius = st->radius;
    }
    row.cumulative_time_in_seconds = lm_ticks(st);
    if (st->trace_len < d.trace_cap) d.trace[st->trace_len] = row;
    st->trace_len++;
    if (st->iteration >= o.max_iterations)
        lm_finish(st, BA_NO_CONVERGENCE, BA_REASON_MAX_ITERATIONS, st->iteration);
    else if (!st->accepted && st->gmax <= o.gtol)
        lm_finish(st, BA_CONVERGENCE, BA_REASON_GRADIENT_TOLERANCE, st->gmax);
    else if (!st->accepted && st->radius <= o.min_radius)
        lm_finish(st, BA_CONVERGENCE, BA_REASON_MIN_RADIUS, st->radius);
}

// Accepted step: x <- candidate (cameras then points, one thread per parameter).
template <typename Real>
__global__ void __launch_bounds__(kWG) k_apply(Dev<Real> d) {
    if (!d.st->accepted) return;
    const size_t i = (size_t)blockIdx.x * kWG + threadIdx.x;
    const size_t ncam = 9 * (size_t)d.nc, npt = 4 * (size_t)d.np;
    if (i < ncam) d.cam_x[i] = d.cam_c[i];
    else if (i < ncam + npt) d.pt_x[i - ncam] = d.pt_c[i - ncam];
}

}  // namespace mba
