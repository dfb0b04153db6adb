// lm_kernels.h -- device kernels of the Levenberg-Marquardt bundle-adjustment solve.
//
// Device layout (built once per problem by the host, see solver.hip):
//   * observations in camera order (CSR by camera, cam_off): co_pt (point), co_uv (x,y);
//     every Jacobian is recomputed in camera-order kernels, where the camera -- and so its
//     CameraPrep -- is uniform across the workgroup, from 12-20 bytes of observation data plus a
//     gathered point.  No per-observation Jacobian is stored.
//   * points sorted by track length (descending) in chunks of 64 (one wave): slot_j holds, for
//     row r of chunk ch and lane l, the camera-order index of that point's r-th observation
//     (-1 padding) at chunk_off[ch] + 64 r + l.  Point kernels sum per-observation partials
//     written by camera kernels, coalesced and load-balanced inside a wave.
//   * parameters FP64 SoA: cam_x[i * nc + c], pt_x[i * np + p].
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

namespace mba {

constexpr int kWG = 256;    // threads per workgroup for every solver kernel
constexpr int kWaves = kWG / 64;

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
    double pcg_rz, pcg_alpha, pcg_beta, pcg_bnorm;
    int iteration, consecutive_invalid, done, termination;
    int need_lin, first_lin, accepted, step_invalid;
    int num_successful, num_unsuccessful;
    int pcg_iters, pcg_done, total_pcg_iters, trace_len;
};

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
};

template <typename Real>
struct Dev {
    int nc, np, no, n_chunks, n_pt_blocks;
    Opts o;
    const int* cam_off;
    const int* co_pt;
    const double* co_uv64;
    const Real* co_uv;
    const int* chunk_off;
    const int* chunk_len;
    const int* slot_j;
    double *cam_x, *pt_x, *cam_c, *pt_c, *dcam, *dpt;
    CameraPrep<Real>* prep;
    CameraPrep<double>* prep_c;
    double *U, *gc, *sc;       // per camera: 81, 9, 9
    double *Ustar, *M, *Minv;  // per camera: 81 each
    double *bc, *px, *pr, *pz, *pp;  // per camera: 9 each (PCG vectors, FP64)
    double *V, *gp, *sp;       // per point: 6, 3, 3
    double* Vinv64;            // per point: 6 (FP64 inverse of the damped block)
    Real* VinvR;               // per point: 6 (same, in working precision)
    Real* qp;                  // per point: 3
    Real* zs;                  // per point: 3
    Real* pv;                  // per observation (camera order): 9 (J_p'J_p sym 6, J_p'r 3)
    Real* u;                   // per observation (camera order): 3
    double* cam_part;          // PS_COUNT * nc
    double* pt_part;           // PS_COUNT * n_pt_blocks
    LMState* st;
    ba_iteration* trace;
    int trace_cap;
};

// ---------------------------------------------------------------------------------------------
// Reductions

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const T o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Sums N per-thread values over the 256-thread workgroup. On return lds[i * kWaves] holds the
// total of value i for every thread (a barrier is included).
template <int N, typename T>
__device__ __forceinline__ void block_sum(T (&v)[N], T* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const T s = wave_sum(v[i]);
        if (lane == 0) lds[i * kWaves + wave] = s;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += kWG) {
        T s = lds[i * kWaves];
        for (int w = 1; w < kWaves; ++w) s += lds[i * kWaves + w];
        lds[i * kWaves] = s;
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
__device__ __forceinline__ void load_point(const double* __restrict__ pt_x, int np, int p, Real X[3]) {
    X[0] = (Real)pt_x[p];
    X[1] = (Real)pt_x[np + p];
    X[2] = (Real)pt_x[2 * np + p];
}

// Corrected (sqrt(rho')-weighted) residual and Jacobians of observation j of camera prep P.
template <typename Real>
__device__ __forceinline__ void corrected_jacobian(const Dev<Real>& d, const CameraPrep<Real>& P, int j, int& p,
                                                   Real r[2], Real Jc[18], Real Jp[6]) {
    p = d.co_pt[j];
    double X[3];
    load_point<double>(d.pt_x, d.np, p, X);
    snavely_residual_jacobian<Real>(P, X, d.co_uv[2 * (size_t)j], d.co_uv[2 * (size_t)j + 1], r, Jc, Jp);
    const Real w = huber_sqrt_rho1<Real>(r[0] * r[0] + r[1] * r[1], d.o.loss, (Real)d.o.loss_a);
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

// One workgroup per camera: U_c = sum Jc'Jc, g_c = sum Jc'r (unscaled, corrected), plus the point
// contributions Jp'Jp and Jp'r of each observation for k_lin_pt.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_lin_cam(Dev<Real> d) {
    if (!d.st->need_lin) return;
    __shared__ double lds[54 * kWaves];
    const int c = blockIdx.x;
    const CameraPrep<Real> P = d.prep[c];
    double acc[54];
#pragma unroll
    for (int i = 0; i < 54; ++i) acc[i] = 0.0;
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6];
        corrected_jacobian<Real>(d, P, j, p, r, Jc, Jp);
        Real hu[45];
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int k = i; k < 9; ++k) hu[sym9(i, k)] = Jc[i] * Jc[k] + Jc[9 + i] * Jc[9 + k];
#pragma unroll
        for (int i = 0; i < 45; ++i) acc[i] += (double)hu[i];
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[45 + i] += (double)(Jc[i] * r[0] + Jc[9 + i] * r[1]);
        Real* pv = d.pv + 9 * (size_t)j;
        pv[0] = Jp[0] * Jp[0] + Jp[3] * Jp[3];
        pv[1] = Jp[0] * Jp[1] + Jp[3] * Jp[4];
        pv[2] = Jp[0] * Jp[2] + Jp[3] * Jp[5];
        pv[3] = Jp[1] * Jp[1] + Jp[4] * Jp[4];
        pv[4] = Jp[1] * Jp[2] + Jp[4] * Jp[5];
        pv[5] = Jp[2] * Jp[2] + Jp[5] * Jp[5];
        pv[6] = Jp[0] * r[0] + Jp[3] * r[1];
        pv[7] = Jp[1] * r[0] + Jp[4] * r[1];
        pv[8] = Jp[2] * r[0] + Jp[5] * r[1];
    }
    block_sum<54>(acc, lds);
    const int t = threadIdx.x;
    if (t < 81) {
        const int i = t / 9, k = t % 9;
        d.U[81 * (size_t)c + t] = lds[sym9(i < k ? i : k, i < k ? k : i) * kWaves];
    }
    if (t < 9) d.gc[9 * (size_t)c + t] = lds[(45 + t) * kWaves];
    if (t == 0) {
        double gm = 0.0;
        for (int i = 0; i < 9; ++i) gm = fmax(gm, fabs(lds[(45 + i) * kWaves]));
        d.cam_part[PS_GMAX * (size_t)d.nc + c] = gm;
        if (d.st->first_lin)
            for (int i = 0; i < 9; ++i)
                d.sc[9 * (size_t)c + i] = d.o.jacobi_scaling ? 1.0 / (1.0 + sqrt(lds[sym9(i, i) * kWaves])) : 1.0;
    }
}

// One thread per point (SELL chunks): V_p = sum Jp'Jp, g_p = sum Jp'r.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_lin_pt(Dev<Real> d) {
    if (!d.st->need_lin) return;
    __shared__ double lds[kWaves];
    const int gid = blockIdx.x * kWG + threadIdx.x;
    const int ch = gid >> 6, lane = gid & 63;
    double gm = 0.0;
    if (ch < d.n_chunks) {
        double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int len = d.chunk_len[ch], base = d.chunk_off[ch];
        for (int r = 0; r < len; ++r) {
            const int j = d.slot_j[base + 64 * r + lane];
            if (j >= 0) {
                const Real* pv = d.pv + 9 * (size_t)j;
#pragma unroll
                for (int i = 0; i < 9; ++i) a[i] += (double)pv[i];
            }
        }
        const int p = gid;
        if (p < d.np) {
            for (int i = 0; i < 6; ++i) d.V[6 * (size_t)p + i] = a[i];
            for (int i = 0; i < 3; ++i) d.gp[3 * (size_t)p + i] = a[6 + i];
            gm = fmax(fabs(a[6]), fmax(fabs(a[7]), fabs(a[8])));
            if (d.st->first_lin) {
                const double dg[3] = {a[0], a[3], a[5]};
                for (int i = 0; i < 3; ++i) d.sp[3 * (size_t)p + i] = d.o.jacobi_scaling ? 1.0 / (1.0 + sqrt(dg[i])) : 1.0;
            }
        }
    }
    gm = block_max1(gm, lds);
    if (threadIdx.x == 0) d.pt_part[PS_GMAX * (size_t)d.n_pt_blocks + blockIdx.x] = gm;
}

// FP64 cost 1/2 sum rho at the parameters whose prep is prep_c and points pts (one WG per camera).
template <typename Real>
__global__ void __launch_bounds__(kWG) k_cost_cam(Dev<Real> d, const double* __restrict__ pts) {
    __shared__ double lds[kWaves];
    const int c = blockIdx.x;
    const CameraPrep<double> P = d.prep_c[c];
    double acc = 0.0;
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        const int p = d.co_pt[j];
        double X[3];
        load_point<double>(pts, d.np, p, X);
        double r[2];
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
    if (threadIdx.x != 0) return;
    st->gmax = gm;
    if (st->first_lin) {
        st->x_cost = cost;
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
        if (!isfinite(cost)) {
            st->done = 1;
            st->termination = BA_FAILURE;
        } else if (st->iteration >= d.o.max_iterations) {
            st->done = 1;
            st->termination = BA_NO_CONVERGENCE;
        }
    } else if (st->trace_len > 0 && st->trace_len <= d.trace_cap) {
        d.trace[st->trace_len - 1].gradient_max_norm = gm;
    }
    if (!st->done && gm <= d.o.gtol) {
        st->done = 1;
        st->termination = BA_CONVERGENCE;
    }
    if (!st->done && st->radius <= d.o.min_radius) {
        st->done = 1;
        st->termination = BA_CONVERGENCE;
    }
    st->need_lin = 0;
    st->first_lin = 0;
}

// ---------------------------------------------------------------------------------------------
// Damped blocks and the reduced camera system

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
    for (int i = 0; i < 3; ++i) A[4 * i] += fmin(fmax(A[4 * i], d.o.min_diag), d.o.max_diag) / radius;
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
    for (int i = 0; i < 3; ++i) d.qp[3 * (size_t)p + i] = (Real)(Ai[3 * i] * g[0] + Ai[3 * i + 1] * g[1] + Ai[3 * i + 2] * g[2]);
}

template <typename Real>
__device__ __forceinline__ void sym3_mul(const Real* S6, const Real v[3], Real out[3]) {
    out[0] = S6[0] * v[0] + S6[1] * v[1] + S6[2] * v[2];
    out[1] = S6[1] * v[0] + S6[3] * v[1] + S6[4] * v[2];
    out[2] = S6[2] * v[0] + S6[4] * v[1] + S6[5] * v[2];
}

// One WG per camera: b_c = S g_c - sum W~ q_p, the Schur-Jacobi block
// M_c = U*_c - sum W~ Vinv W~' (U*_c = S U S + D_c^2).  With W~ = S_c Jc' Jp S_p, a point's
// contribution is Jc~' (Jp~ Vinv Jp~') Jc~, a 2x2 middle factor per observation.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_rhs_cam(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done) return;
    __shared__ double lds[54 * kWaves];
    const int c = blockIdx.x;
    const CameraPrep<Real> P = d.prep[c];
    Real sc[9];
    for (int i = 0; i < 9; ++i) sc[i] = (Real)d.sc[9 * (size_t)c + i];
    double acc[54];
#pragma unroll
    for (int i = 0; i < 54; ++i) acc[i] = 0.0;
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6];
        corrected_jacobian<Real>(d, P, j, p, r, Jc, Jp);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            Jc[i] *= sc[i];
            Jc[9 + i] *= sc[i];
        }
        const Real s0 = (Real)d.sp[3 * (size_t)p], s1 = (Real)d.sp[3 * (size_t)p + 1], s2 = (Real)d.sp[3 * (size_t)p + 2];
        Jp[0] *= s0; Jp[3] *= s0;
        Jp[1] *= s1; Jp[4] *= s1;
        Jp[2] *= s2; Jp[5] *= s2;
        const Real* q = d.qp + 3 * (size_t)p;
        const Real a0 = Jp[0] * q[0] + Jp[1] * q[1] + Jp[2] * q[2];
        const Real a1 = Jp[3] * q[0] + Jp[4] * q[1] + Jp[5] * q[2];
        const Real* Vi = d.VinvR + 6 * (size_t)p;
        Real h0[3], h1[3];
        sym3_mul<Real>(Vi, Jp, h0);
        sym3_mul<Real>(Vi, Jp + 3, h1);
        const Real G00 = Jp[0] * h0[0] + Jp[1] * h0[1] + Jp[2] * h0[2];
        const Real G01 = Jp[0] * h1[0] + Jp[1] * h1[1] + Jp[2] * h1[2];
        const Real G11 = Jp[3] * h1[0] + Jp[4] * h1[1] + Jp[5] * h1[2];
        Real H0[9], H1[9];  // G * Jc~ (2x9)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            H0[k] = G00 * Jc[k] + G01 * Jc[9 + k];
            H1[k] = G01 * Jc[k] + G11 * Jc[9 + k];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int k = i; k < 9; ++k) acc[sym9(i, k)] += (double)(Jc[i] * H0[k] + Jc[9 + i] * H1[k]);
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[45 + i] += (double)(Jc[i] * a0 + Jc[9 + i] * a1);
    }
    block_sum<54>(acc, lds);
    const int t = threadIdx.x;
    if (t < 81) {
        const int i = t / 9, k = t % 9;
        const double si = d.sc[9 * (size_t)c + i], sk = d.sc[9 * (size_t)c + k];
        double u = si * d.U[81 * (size_t)c + t] * sk;
        if (i == k) u += fmin(fmax(u, d.o.min_diag), d.o.max_diag) / st->radius;
        d.Ustar[81 * (size_t)c + t] = u;
        d.M[81 * (size_t)c + t] = u - lds[sym9(i < k ? i : k, i < k ? k : i) * kWaves];
    }
    if (t < 9) d.bc[9 * (size_t)c + t] = d.sc[9 * (size_t)c + t] * d.gc[9 * (size_t)c + t] - lds[(45 + t) * kWaves];
}

// One thread per camera: Minv_c = M_c^-1 (FP64) and the PCG start x = 0, r = b, z = Minv b, p = 0.
template <typename Real>
__global__ void __launch_bounds__(64) k_cam_factor(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.nc) return;
    double A[81];
    for (int i = 0; i < 81; ++i) A[i] = d.M[81 * (size_t)c + i];
    double rz = 0.0, bb = 0.0;
    if (!chol_inplace<9>(A)) {
        st->step_invalid = 1;
    } else {
        double Ai[81];
        chol_inverse<9>(A, Ai);
        double b[9];
        for (int i = 0; i < 9; ++i) b[i] = d.bc[9 * (size_t)c + i];
        for (int i = 0; i < 9; ++i) {
            double z = 0.0;
            for (int k = 0; k < 9; ++k) z += Ai[9 * i + k] * b[k];
            d.pz[9 * (size_t)c + i] = z;
            d.pr[9 * (size_t)c + i] = b[i];
            d.px[9 * (size_t)c + i] = 0.0;
            d.pp[9 * (size_t)c + i] = 0.0;
            rz += b[i] * z;
            bb += b[i] * b[i];
        }
        for (int i = 0; i < 81; ++i) d.Minv[81 * (size_t)c + i] = Ai[i];
    }
    d.cam_part[PS_A * (size_t)d.nc + c] = rz;
    d.cam_part[PS_B * (size_t)d.nc + c] = bb;
}

// ---------------------------------------------------------------------------------------------
// Implicit-Schur PCG.  Per iteration: k_pcg_beta -> k_jtx(update p) -> k_pt_z -> k_pcg_alpha -> k_jz.

// Single WG. first != 0: rz = sum r.z, |b| = sqrt(sum b.b), p = z (beta 0). Otherwise check
// ||r|| <= tol ||b|| (partials of k_jz), count the finished iteration, beta = rz_new / rz.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_pcg_beta(Dev<Real> d, int first) {
    __shared__ double lds[2 * kWaves];
    LMState* st = d.st;
    if (st->done || st->step_invalid || (!first && st->pcg_done)) return;
    double a[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < d.nc; i += kWG) {
        a[0] += d.cam_part[PS_A * (size_t)d.nc + i];
        a[1] += d.cam_part[PS_B * (size_t)d.nc + i];
    }
    block_sum<2>(a, lds);
    if (threadIdx.x != 0) return;
    const double rz = lds[0], rr = lds[kWaves];
    if (first) {
        st->pcg_iters = 0;
        st->pcg_done = 0;
        st->pcg_bnorm = sqrt(rr);
        st->pcg_rz = rz;
        st->pcg_beta = 0.0;
        if (!(st->pcg_bnorm > 0.0)) st->pcg_done = 1;
        if (!isfinite(rz) || !isfinite(rr)) st->step_invalid = 1;
        if (d.o.pcg_max_iterations <= 0) st->pcg_done = 1;
        return;
    }
    st->pcg_iters++;
    if (sqrt(rr) <= d.o.pcg_tol * st->pcg_bnorm || st->pcg_iters >= d.o.pcg_max_iterations) {
        st->pcg_done = 1;
        return;
    }
    st->pcg_beta = rz / st->pcg_rz;
    st->pcg_rz = rz;
}

// One WG per camera.  v = p (after p <- z + beta p when update != 0) or v = y (back-substitution):
// u_j = Jp' (Jc (S_c v_c)) per observation; partial v'U*v for the PCG step length.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_jtx(Dev<Real> d, int update, const double* __restrict__ vsrc) {
    LMState* st = d.st;
    if (st->done || st->step_invalid || (update && st->pcg_done)) return;
    __shared__ double vsh[9];
    const int c = blockIdx.x;
    if (threadIdx.x < 9) {
        const size_t i = 9 * (size_t)c + threadIdx.x;
        double v = vsrc[i];
        if (update) {
            v = d.pz[i] + st->pcg_beta * v;
            d.pp[i] = v;
        }
        vsh[threadIdx.x] = v;
    }
    __syncthreads();
    if (update && threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < 9; ++i) {
            double t = 0.0;
            for (int k = 0; k < 9; ++k) t += d.Ustar[81 * (size_t)c + 9 * i + k] * vsh[k];
            s += vsh[i] * t;
        }
        d.cam_part[PS_A * (size_t)d.nc + c] = s;
    }
    Real vs[9];
    for (int i = 0; i < 9; ++i) vs[i] = (Real)(d.sc[9 * (size_t)c + i] * vsh[i]);
    const CameraPrep<Real> P = d.prep[c];
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6];
        corrected_jacobian<Real>(d, P, j, p, r, Jc, Jp);
        Real a0 = 0, a1 = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            a0 += Jc[i] * vs[i];
            a1 += Jc[9 + i] * vs[i];
        }
        Real* u = d.u + 3 * (size_t)j;
        u[0] = Jp[0] * a0 + Jp[3] * a1;
        u[1] = Jp[1] * a0 + Jp[4] * a1;
        u[2] = Jp[2] * a0 + Jp[5] * a1;
    }
}

// One thread per point: t_p = S_p sum u_j, z_p = Vinv t_p, zs_p = S_p z_p; partial t.z.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_pt_z(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid || st->pcg_done) return;
    __shared__ double lds[kWaves];
    const int gid = blockIdx.x * kWG + threadIdx.x;
    const int ch = gid >> 6, lane = gid & 63;
    double tz = 0.0;
    if (ch < d.n_chunks && gid < d.np) {
        Real a[3] = {0, 0, 0};
        const int len = d.chunk_len[ch], base = d.chunk_off[ch];
        for (int r = 0; r < len; ++r) {
            const int j = d.slot_j[base + 64 * r + lane];
            if (j >= 0) {
                const Real* u = d.u + 3 * (size_t)j;
                a[0] += u[0];
                a[1] += u[1];
                a[2] += u[2];
            }
        }
        const int p = gid;
        Real s[3], t[3], z[3];
        for (int i = 0; i < 3; ++i) {
            s[i] = (Real)d.sp[3 * (size_t)p + i];
            t[i] = s[i] * a[i];
        }
        sym3_mul<Real>(d.VinvR + 6 * (size_t)p, t, z);
        for (int i = 0; i < 3; ++i) d.zs[3 * (size_t)p + i] = s[i] * z[i];
        tz = (double)t[0] * (double)z[0] + (double)t[1] * (double)z[1] + (double)t[2] * (double)z[2];
    }
    tz = block_sum1(tz, lds);
    if (threadIdx.x == 0) d.pt_part[PS_B * (size_t)d.n_pt_blocks + blockIdx.x] = tz;
}

// Single WG: alpha = rz / (p'Sp), p'Sp = sum_c p_c'U*_c p_c - sum_p t_p'Vinv t_p.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_pcg_alpha(Dev<Real> d) {
    __shared__ double lds[2 * kWaves];
    LMState* st = d.st;
    if (st->done || st->step_invalid || st->pcg_done) return;
    double a[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < d.nc; i += kWG) a[0] += d.cam_part[PS_A * (size_t)d.nc + i];
    for (int i = threadIdx.x; i < d.n_pt_blocks; i += kWG) a[1] += d.pt_part[PS_B * (size_t)d.n_pt_blocks + i];
    block_sum<2>(a, lds);
    if (threadIdx.x != 0) return;
    const double pq = lds[0] - lds[kWaves];
    if (!(pq > 0.0) || !isfinite(pq)) {
        // S is positive definite in exact arithmetic; a non-positive curvature means the search
        // direction is at rounding level: stop with the current iterate.
        st->pcg_done = 1;
        return;
    }
    st->pcg_alpha = st->pcg_rz / pq;
}

// One WG per camera: q_c = U*_c p_c - S_c sum Jc'(Jp zs_p); x += alpha p, r -= alpha q,
// z = Minv r; partials r.z and r.r.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_jz(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid || st->pcg_done) return;
    __shared__ double lds[9 * kWaves];
    __shared__ double rsh[9];
    const int c = blockIdx.x;
    const CameraPrep<Real> P = d.prep[c];
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = 0.0;
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6];
        corrected_jacobian<Real>(d, P, j, p, r, Jc, Jp);
        const Real* z = d.zs + 3 * (size_t)p;
        const Real a0 = Jp[0] * z[0] + Jp[1] * z[1] + Jp[2] * z[2];
        const Real a1 = Jp[3] * z[0] + Jp[4] * z[1] + Jp[5] * z[2];
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += (double)(Jc[i] * a0 + Jc[9 + i] * a1);
    }
    block_sum<9>(acc, lds);
    const int t = threadIdx.x;
    const double alpha = st->pcg_alpha;
    if (t < 9) {
        const size_t i = 9 * (size_t)c + t;
        double uq = 0.0;
        for (int k = 0; k < 9; ++k) uq += d.Ustar[81 * (size_t)c + 9 * t + k] * d.pp[9 * (size_t)c + k];
        const double q = uq - d.sc[i] * lds[t * kWaves];
        d.px[i] += alpha * d.pp[i];
        const double r = d.pr[i] - alpha * q;
        d.pr[i] = r;
        rsh[t] = r;
    }
    __syncthreads();
    if (t == 0) {
        double rz = 0.0, rr = 0.0;
        for (int i = 0; i < 9; ++i) {
            double z = 0.0;
            for (int k = 0; k < 9; ++k) z += d.Minv[81 * (size_t)c + 9 * i + k] * rsh[k];
            d.pz[9 * (size_t)c + i] = z;
            rz += rsh[i] * z;
            rr += rsh[i] * rsh[i];
        }
        d.cam_part[PS_A * (size_t)d.nc + c] = rz;
        d.cam_part[PS_B * (size_t)d.nc + c] = rr;
    }
}

// ---------------------------------------------------------------------------------------------
// Step, candidate and decision

// One thread per point: y_p = Vinv (S g_p - S sum u_j), delta_p = -S y_p, candidate = x + delta.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_pt_backsub(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid) return;
    __shared__ double lds[2 * kWaves];
    const int gid = blockIdx.x * kWG + threadIdx.x;
    const int ch = gid >> 6, lane = gid & 63;
    double a2[2] = {0.0, 0.0};
    if (ch < d.n_chunks && gid < d.np) {
        double a[3] = {0, 0, 0};
        const int len = d.chunk_len[ch], base = d.chunk_off[ch];
        for (int r = 0; r < len; ++r) {
            const int j = d.slot_j[base + 64 * r + lane];
            if (j >= 0) {
                const Real* u = d.u + 3 * (size_t)j;
                a[0] += (double)u[0];
                a[1] += (double)u[1];
                a[2] += (double)u[2];
            }
        }
        const int p = gid;
        double s[3], t[3];
        for (int i = 0; i < 3; ++i) {
            s[i] = d.sp[3 * (size_t)p + i];
            t[i] = s[i] * d.gp[3 * (size_t)p + i] - s[i] * a[i];
        }
        const double* Vi = d.Vinv64 + 6 * (size_t)p;
        const double y[3] = {Vi[0] * t[0] + Vi[1] * t[1] + Vi[2] * t[2], Vi[1] * t[0] + Vi[3] * t[1] + Vi[4] * t[2],
                             Vi[2] * t[0] + Vi[4] * t[1] + Vi[5] * t[2]};
        for (int i = 0; i < 3; ++i) {
            const double dl = -s[i] * y[i];
            const double x = d.pt_x[(size_t)i * d.np + p];
            d.dpt[(size_t)i * d.np + p] = dl;
            d.pt_c[(size_t)i * d.np + p] = x + dl;
            a2[0] += dl * dl;
            a2[1] += x * x;
        }
    }
    block_sum<2>(a2, lds);
    if (threadIdx.x == 0) {
        d.pt_part[PS_DX2 * (size_t)d.n_pt_blocks + blockIdx.x] = lds[0];
        d.pt_part[PS_X2 * (size_t)d.n_pt_blocks + blockIdx.x] = lds[kWaves];
    }
}

// One thread per camera: delta_c = -S_c y_c, candidate camera and its FP64 prep.
template <typename Real>
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
        d.cam_c[(size_t)i * d.nc + c] = cam[i];
        dx2 += dl * dl;
        x2 += x * x;
    }
    CameraPrep<double> P;
    camera_prep<double>(cam, P);
    d.prep_c[c] = P;
    d.cam_part[PS_DX2 * (size_t)d.nc + c] = dx2;
    d.cam_part[PS_X2 * (size_t)d.nc + c] = x2;
}

// One WG per camera: model cost change -(J~ s).(r~ + J~ s / 2) with J, r at x (working
// precision), and the FP64 candidate cost at x + delta.
template <typename Real>
__global__ void __launch_bounds__(kWG) k_model_cand(Dev<Real> d) {
    LMState* st = d.st;
    if (st->done || st->step_invalid) return;
    __shared__ double lds[2 * kWaves];
    const int c = blockIdx.x;
    const CameraPrep<Real> P = d.prep[c];
    const CameraPrep<double> Pc = d.prep_c[c];
    Real dc[9];
    for (int i = 0; i < 9; ++i) dc[i] = (Real)d.dcam[(size_t)i * d.nc + c];
    double acc[2] = {0.0, 0.0};
    for (int j = d.cam_off[c] + threadIdx.x; j < d.cam_off[c + 1]; j += kWG) {
        int p;
        Real r[2], Jc[18], Jp[6];
        corrected_jacobian<Real>(d, P, j, p, r, Jc, Jp);
        const Real dp[3] = {(Real)d.dpt[p], (Real)d.dpt[(size_t)d.np + p], (Real)d.dpt[2 * (size_t)d.np + p]};
        Real m0 = Jp[0] * dp[0] + Jp[1] * dp[1] + Jp[2] * dp[2];
        Real m1 = Jp[3] * dp[0] + Jp[4] * dp[1] + Jp[5] * dp[2];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            m0 += Jc[i] * dc[i];
            m1 += Jc[9 + i] * dc[i];
        }
        acc[0] += (double)m0 * ((double)r[0] + 0.5 * (double)m0) + (double)m1 * ((double)r[1] + 0.5 * (double)m1);
        double X[3];
        load_point<double>(d.pt_c, d.np, p, X);
        double rc[2];
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
    const bool valid = have_step && model > 0.0 && isfinite(dx2);
    if (!valid) {
        st->step_invalid = 0;
        if (++st->consecutive_invalid >= o.max_invalid) {
            st->done = 1;
            st->termination = BA_FAILURE;
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
            st->done = 1;
            st->termination = BA_CONVERGENCE;
            return;
        }
        const double cost_change = st->x_cost - cand;
        row.cost_change = cost_change;
        if (fabs(cost_change) <= o.ftol * st->x_cost) {
            st->done = 1;
            st->termination = BA_CONVERGENCE;
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
    if (st->trace_len < d.trace_cap) d.trace[st->trace_len] = row;
    st->trace_len++;
    if (st->iteration >= o.max_iterations) {
        st->done = 1;
        st->termination = BA_NO_CONVERGENCE;
    } else if (!st->accepted && st->gmax <= o.gtol) {
        st->done = 1;
        st->termination = BA_CONVERGENCE;
    } else if (!st->accepted && st->radius <= o.min_radius) {
        st->done = 1;
        st->termination = BA_CONVERGENCE;
    }
}

// Accepted step: x <- candidate (cameras then points, one thread per parameter).
template <typename Real>
__global__ void __launch_bounds__(kWG) k_apply(Dev<Real> d) {
    if (!d.st->accepted) return;
    const size_t i = (size_t)blockIdx.x * kWG + threadIdx.x;
    const size_t ncam = 9 * (size_t)d.nc, npt = 3 * (size_t)d.np;
    if (i < ncam) d.cam_x[i] = d.cam_c[i];
    else if (i < ncam + npt) d.pt_x[i - ncam] = d.pt_c[i - ncam];
}

}  // namespace mba
