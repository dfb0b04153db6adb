// window_kernels.h -- batched local bundle adjustment: one workgroup runs whole LM iterations of
// one window (the dense-Schur solve of ba_solve, restated for windows of <= 32 cameras).
//
// Per LM iteration, with __syncthreads between phases:
//   lin    (after an accepted step) camera prep; waves over cameras recompute the corrected
//          Jacobians: diag(Jc'Jc), Jc'r per camera by wave reductions, Jp'Jp and Jp'r per
//          observation (FP64 products) into slot order; threads over points sum their slots;
//   damp   V*^-1 (FP64), q = V*^-1 g~_p, D_c^2;
//   store  per observation Jc~, Jp~, Hp = Jp~ V*^-1, I - G (FP64-formed), r~; reduced gradient
//          b_c = sum Jc~'(r~ - Jp~ q) by wave reductions;
//   S      each wave owns 9x9 blocks (a, b) of S and reduces their observation pairs in a fixed
//          order, staging 16 pairs at a time in wave-private LDS (no workgroup barrier);
//   chol   S in LDS, tiled right-looking Cholesky (16-column panels factored by wave 0, trailing
//          update by the workgroup: 2 barriers per panel); triangular solves by one wave with
//          the right-hand side in registers and shuffles; FP64 refinement against the FP64 S in
//          mixed precision;
//   back   u = Jp~'(r~ - Jc~ y_c) per observation, y_p = V*^-1 sum u per point, candidate;
//   decide model cost change from the stored J~, FP64 candidate cost, Ceres LM rule.
// A window's arithmetic and reduction order depend only on the window, so its result is the same
// whichever other windows share the launch.
#pragma once

#include "window.h"

namespace mba {

__device__ __forceinline__ void wave_sync() {
    // LDS traffic of one wave is processed in order; this keeps the compiler from moving loads
    // and stores of the wave-private staging area across the exchange point
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum of N values over the workgroup (NT threads); totals in red[i * NW], barrier included.
template <int N, int NT>
__device__ __forceinline__ void wg_sum(double (&v)[N], double* red) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double s = wave_sum(v[i]);
        if (lane == 0) red[i * NW + wave] = s;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        double s = red[threadIdx.x * NW];
        for (int w = 1; w < NW; ++w) s += red[threadIdx.x * NW + w];
        red[threadIdx.x * NW] = s;
    }
    __syncthreads();
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

struct WinPtrs {
    const int *cam_off, *co_pt, *slot_of, *chunk_len, *chunk_off;
    const int4* blk;
    const int2* pairs;
    const double* co_uv64;
    double *cam_x, *cam_c, *udiag, *gc, *sc, *dc2, *bc, *yc, *V, *gp, *sp, *Vinv, *yp, *S64;
    double4 *pt_x, *pt_c, *pv;
    WinState* st;
};

template <typename Real>
__device__ __forceinline__ void win_corrected(const CameraPrep<Real>& P, const double4* pt_x, const Real* co_uv,
                                              const int* co_pt, int j, const Opts& o, int& p, Real r[2], Real Jc[18],
                                              Real Jp[6]) {
    p = co_pt[j];
    const double4 x4 = pt_x[p];
    const double X[3] = {x4.x, x4.y, x4.z};
    snavely_residual_jacobian<Real>(P, X, co_uv[2 * (size_t)j], co_uv[2 * (size_t)j + 1], r, Jc, Jp);
    const Real w = huber_sqrt_rho1<Real>(r[0] * r[0] + r[1] * r[1], o.loss, (Real)o.loss_a);
    r[0] *= w;
    r[1] *= w;
#pragma unroll
    for (int i = 0; i < 18; ++i) Jc[i] *= w;
#pragma unroll
    for (int i = 0; i < 6; ++i) Jp[i] *= w;
}

// Forward (lower = 1: L z = b) or backward (lower = 0: L' x = z) substitution of one wave with the
// vector held as yv[s] = y[64 s + lane]; L is the n x n lower factor in LDS.
template <typename F, int NS>
__device__ __forceinline__ void wave_trsv(const F* L, int n, double (&yv)[NS], bool lower) {
    const int lane = threadIdx.x & 63;
    if (lower) {
#pragma unroll
        for (int ks = 0; ks < NS; ++ks) {
            for (int kk = 0; kk < 64; ++kk) {
                const int k = 64 * ks + kk;
                if (k >= n) break;
                const double v = __shfl(yv[ks], kk, 64) / (double)L[(size_t)k * n + k];
                if (lane == kk) yv[ks] = v;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const int i = 64 * s + lane;
                    if (i > k && i < n) yv[s] -= (double)L[(size_t)i * n + k] * v;
                }
            }
        }
    } else {
#pragma unroll
        for (int ks = NS - 1; ks >= 0; --ks) {
            for (int kk = 63; kk >= 0; --kk) {
                const int k = 64 * ks + kk;
                if (k >= n) continue;
                const double v = __shfl(yv[ks], kk, 64) / (double)L[(size_t)k * n + k];
                if (lane == kk) yv[ks] = v;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const int i = 64 * s + lane;
                    if (i < k) yv[s] -= (double)L[(size_t)k * n + i] * v;
                }
            }
        }
    }
}

constexpr int kWinMaxN = 288;  // 9 * 32 cameras
constexpr int kWinNS = (kWinMaxN + 63) / 64;
constexpr int kWinTile = 16;

template <typename Real, int NT>
__global__ void __launch_bounds__(NT) k_window_lm(const WinDesc* __restrict__ descs, char* __restrict__ arena,
                                                  const int* __restrict__ work, int n_work, int iters, Opts o,
                                                  int refine) {
    using F = Real;  // factor precision: FP32 in mixed precision (refined in FP64), FP64 otherwise
    constexpr int NW = NT / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[NW * 20];
    __shared__ Real stage[NW][16][36];
    __shared__ double rsh[kWinMaxN];
    __shared__ int sflag[2];  // 0: step invalid, 1: accepted
    const int wi = blockIdx.x;
    if (wi >= n_work) return;
    const WinDesc D = descs[work[wi]];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int K = D.K, P = D.P, n = 9 * K;
    WinPtrs w;
    w.cam_off = at<int>(arena, D.o_cam_off);
    w.co_pt = at<int>(arena, D.o_co_pt);
    w.slot_of = at<int>(arena, D.o_slot_of);
    w.chunk_len = at<int>(arena, D.o_chunk_len);
    w.chunk_off = at<int>(arena, D.o_chunk_off);
    w.blk = at<int4>(arena, D.o_blk);
    w.pairs = at<int2>(arena, D.o_pairs);
    w.co_uv64 = at<double>(arena, D.o_co_uv64);
    const Real* co_uv = at<Real>(arena, D.o_co_uv);
    w.cam_x = at<double>(arena, D.o_cam_x);
    w.cam_c = at<double>(arena, D.o_cam_c);
    w.udiag = at<double>(arena, D.o_udiag);
    w.gc = at<double>(arena, D.o_gc);
    w.sc = at<double>(arena, D.o_sc);
    w.dc2 = at<double>(arena, D.o_dc2);
    w.bc = at<double>(arena, D.o_bc);
    w.yc = at<double>(arena, D.o_yc);
    w.V = at<double>(arena, D.o_V);
    w.gp = at<double>(arena, D.o_gp);
    w.sp = at<double>(arena, D.o_sp);
    w.Vinv = at<double>(arena, D.o_Vinv);
    w.yp = at<double>(arena, D.o_yp);
    w.S64 = at<double>(arena, D.o_S64);
    w.pt_x = at<double4>(arena, D.o_pt_x);
    w.pt_c = at<double4>(arena, D.o_pt_c);
    w.pv = at<double4>(arena, D.o_pv);
    w.st = at<WinState>(arena, D.o_state);
    CameraPrep<Real>* prep = at<CameraPrep<Real>>(arena, D.o_prep);
    CameraPrep<double>* prep_c = at<CameraPrep<double>>(arena, D.o_prep_c);
    R4<Real>* qp = at<R4<Real>>(arena, D.o_qp);
    R4<Real>* u = at<R4<Real>>(arena, D.o_u);
    Real* js = at<Real>(arena, D.o_js);
    WinState* st = w.st;
    F* Lf = reinterpret_cast<F*>(smem);

    __syncthreads();
    if (st->done) return;
    if (st->first_lin) {
        // iteration 0: the FP64 cost at the initial parameters
        for (int c = tid; c < K; c += NT) {
            CameraPrep<double> Pc;
            camera_prep<double>(w.cam_x + 9 * (size_t)c, Pc);
            prep_c[c] = Pc;
        }
        __syncthreads();
        double cost[1] = {0.0};
        for (int c = wave; c < K; c += NW) {
            const CameraPrep<double> Pc = prep_c[c];
            for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                const double4 x4 = w.pt_x[w.co_pt[j]];
                const double X[3] = {x4.x, x4.y, x4.z};
                double r[2];
                snavely_residual<double>(Pc, X, w.co_uv64[2 * (size_t)j], w.co_uv64[2 * (size_t)j + 1], r);
                cost[0] += half_rho(r[0] * r[0] + r[1] * r[1], o.loss, o.loss_a);
            }
        }
        wg_sum<1, NT>(cost, red);
        if (tid == 0) {
            st->x_cost = st->initial_cost = red[0];
            st->rows = 1;
            if (!isfinite(red[0])) {
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
                CameraPrep<Real> Pr;
                camera_prep<Real>(w.cam_x + 9 * (size_t)c, Pr);
                prep[c] = Pr;
            }
            __syncthreads();
            double gm = 0.0;
            for (int c = wave; c < K; c += NW) {
                const CameraPrep<Real> Pr = prep[c];
                double acc[18];
#pragma unroll
                for (int i = 0; i < 18; ++i) acc[i] = 0.0;
                for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                    int p;
                    Real r[2], Jc[18], Jp[6];
                    win_corrected<Real>(Pr, w.pt_x, co_uv, w.co_pt, j, o, p, r, Jc, Jp);
#pragma unroll
                    for (int i = 0; i < 9; ++i) {
                        acc[i] += (double)(Jc[i] * Jc[i] + Jc[9 + i] * Jc[9 + i]);
                        acc[9 + i] += (double)(Jc[i] * r[0] + Jc[9 + i] * r[1]);
                    }
                    const double q[6] = {Jp[0], Jp[1], Jp[2], Jp[3], Jp[4], Jp[5]};
                    const double r0 = r[0], r1 = r[1];
                    double4* pv = w.pv + 3 * (size_t)w.slot_of[j];
                    pv[0] = make_double4(q[0] * q[0] + q[3] * q[3], q[0] * q[1] + q[3] * q[4],
                                         q[0] * q[2] + q[3] * q[5], q[1] * q[1] + q[4] * q[4]);
                    pv[1] = make_double4(q[1] * q[2] + q[4] * q[5], q[2] * q[2] + q[5] * q[5], q[0] * r0 + q[3] * r1,
                                         q[1] * r0 + q[4] * r1);
                    pv[2] = make_double4(q[2] * r0 + q[5] * r1, 0.0, 0.0, 0.0);
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
                    const double4* pv = w.pv + 3 * (base + 64 * (size_t)r + l);
                    const double4 x = pv[0], y = pv[1], z = pv[2];
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
        if (tid == 0) sflag[0] = 0;
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
            R4<Real> q;
            q.x = (Real)(Ai[0] * g[0] + Ai[1] * g[1] + Ai[2] * g[2]);
            q.y = (Real)(Ai[3] * g[0] + Ai[4] * g[1] + Ai[5] * g[2]);
            q.z = (Real)(Ai[6] * g[0] + Ai[7] * g[1] + Ai[8] * g[2]);
            q.w = (Real)0;
            qp[p] = q;
        }
        for (int k = tid; k < n; k += NT)
            w.dc2[k] = fmin(fmax(w.sc[k] * w.udiag[k] * w.sc[k], o.min_diag), o.max_diag) / radius;
        __syncthreads();
        const bool invalid_blocks = sflag[0] != 0;
        // ---------------- store + reduced gradient ----------------
        if (!invalid_blocks) {
            for (int c = wave; c < K; c += NW) {
                const CameraPrep<Real> Pr = prep[c];
                Real scc[9];
                for (int i = 0; i < 9; ++i) scc[i] = (Real)w.sc[9 * (size_t)c + i];
                double bacc[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) bacc[i] = 0.0;
                for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                    int p;
                    Real r[2], Jc[18], Jp[6];
                    win_corrected<Real>(Pr, w.pt_x, co_uv, w.co_pt, j, o, p, r, Jc, Jp);
#pragma unroll
                    for (int i = 0; i < 9; ++i) {
                        Jc[i] *= scc[i];
                        Jc[9 + i] *= scc[i];
                    }
                    for (int i = 0; i < 3; ++i) {
                        const Real s = (Real)w.sp[3 * (size_t)p + i];
                        Jp[i] *= s;
                        Jp[3 + i] *= s;
                    }
                    const double* Vi = w.Vinv + 6 * (size_t)p;
                    const double j0[3] = {Jp[0], Jp[1], Jp[2]}, j1[3] = {Jp[3], Jp[4], Jp[5]};
                    double h0[3], h1[3];
                    sym3_mul<double>(Vi, j0, h0);
                    sym3_mul<double>(Vi, j1, h1);
                    Real* rec = js + kWinJs * (size_t)j;
#pragma unroll
                    for (int i = 0; i < 18; ++i) rec[i] = Jc[i];
#pragma unroll
                    for (int i = 0; i < 6; ++i) rec[18 + i] = Jp[i];
                    for (int i = 0; i < 3; ++i) {
                        rec[24 + i] = (Real)h0[i];
                        rec[27 + i] = (Real)h1[i];
                    }
                    rec[30] = (Real)(1.0 - (j0[0] * h0[0] + j0[1] * h0[1] + j0[2] * h0[2]));
                    rec[31] = (Real)(-(j0[0] * h1[0] + j0[1] * h1[1] + j0[2] * h1[2]));
                    rec[32] = (Real)(1.0 - (j1[0] * h1[0] + j1[1] * h1[1] + j1[2] * h1[2]));
                    rec[33] = r[0];
                    rec[34] = r[1];
                    const R4<Real> q = qp[p];
                    const Real e0 = r[0] - (Jp[0] * q.x + Jp[1] * q.y + Jp[2] * q.z);
                    const Real e1 = r[1] - (Jp[3] * q.x + Jp[4] * q.y + Jp[5] * q.z);
#pragma unroll
                    for (int i = 0; i < 9; ++i) bacc[i] += (double)(Jc[i] * e0 + Jc[9 + i] * e1);
                }
#pragma unroll
                for (int i = 0; i < 9; ++i) bacc[i] = wave_sum(bacc[i]);
                if (lane == 0)
                    for (int i = 0; i < 9; ++i) w.bc[9 * (size_t)c + i] = bacc[i];
            }
            __syncthreads();
            // ---------------- S ----------------
            for (int b = wave; b < D.n_blocks; b += NW) {
                const int4 bk = w.blk[b];
                const int ca = bk.x, cb = bk.y;
                const int e0 = lane, e1 = lane + 64;
                const int r0 = e0 / 9, c0 = e0 % 9, r1 = e1 / 9, c1 = e1 % 9;
                double acc0 = 0.0, acc1 = 0.0;
                for (int base = bk.z; base < bk.w; base += 16) {
                    const int m = min(16, bk.w - base);
                    if (lane < m) {
                        const int2 pr = w.pairs[base + lane];
                        const Real* Ji = js + kWinJs * (size_t)pr.x;
                        const Real* Jj = js + kWinJs * (size_t)pr.y;
                        Real G00, G01, G10, G11;
                        if (pr.x == pr.y) {
                            // an observation with itself carries its share of U*: J~c'(I - G)J~c,
                            // negated because the accumulated sum is subtracted
                            G00 = -Ji[30];
                            G01 = G10 = -Ji[31];
                            G11 = -Ji[32];
                        } else {
                            const Real* Hp = Ji + 24;
                            const Real* Jp = Jj + 18;
                            G00 = Hp[0] * Jp[0] + Hp[1] * Jp[1] + Hp[2] * Jp[2];
                            G01 = Hp[0] * Jp[3] + Hp[1] * Jp[4] + Hp[2] * Jp[5];
                            G10 = Hp[3] * Jp[0] + Hp[4] * Jp[1] + Hp[5] * Jp[2];
                            G11 = Hp[3] * Jp[3] + Hp[4] * Jp[4] + Hp[5] * Jp[5];
                        }
                        for (int k = 0; k < 9; ++k) {
                            stage[wave][lane][k] = Ji[k];
                            stage[wave][lane][9 + k] = Ji[9 + k];
                            stage[wave][lane][18 + k] = G00 * Jj[k] + G01 * Jj[9 + k];
                            stage[wave][lane][27 + k] = G10 * Jj[k] + G11 * Jj[9 + k];
                        }
                    }
                    wave_sync();
                    for (int k = 0; k < m; ++k) {
                        const Real* A = stage[wave][k];
                        acc0 += (double)(A[r0] * A[18 + c0] + A[9 + r0] * A[27 + c0]);
                        if (e1 < 81) acc1 += (double)(A[r1] * A[18 + c1] + A[9 + r1] * A[27 + c1]);
                    }
                    wave_sync();
                }
                {
                    const double v = (ca == cb && r0 == c0 ? w.dc2[9 * (size_t)ca + r0] : 0.0) - acc0;
                    w.S64[(size_t)(9 * ca + r0) * n + 9 * cb + c0] = v;
                    if (ca != cb) w.S64[(size_t)(9 * cb + c0) * n + 9 * ca + r0] = v;
                }
                if (e1 < 81) {
                    const double v = (ca == cb && r1 == c1 ? w.dc2[9 * (size_t)ca + r1] : 0.0) - acc1;
                    w.S64[(size_t)(9 * ca + r1) * n + 9 * cb + c1] = v;
                    if (ca != cb) w.S64[(size_t)(9 * cb + c1) * n + 9 * ca + r1] = v;
                }
            }
            __syncthreads();
            // ---------------- chol ----------------
            for (int idx = tid; idx < n * n; idx += NT) Lf[idx] = (F)w.S64[idx];
            if (tid == 0) sflag[1] = 0;
            __syncthreads();
            for (int k0 = 0; k0 < n; k0 += kWinTile) {
                const int kb = min(kWinTile, n - k0);
                if (wave == 0) {
                    for (int j = k0; j < k0 + kb; ++j) {
                        F dj = Lf[(size_t)j * n + j];
                        if (!(dj > (F)0)) {
                            if (lane == 0) sflag[1] = 1;
                            dj = (F)1;
                        }
                        dj = sqrt(dj);
                        wave_sync();
                        if (lane == 0) Lf[(size_t)j * n + j] = dj;
                        for (int i = j + 1 + lane; i < n; i += 64) Lf[(size_t)i * n + j] /= dj;
                        wave_sync();
                        for (int c = j + 1; c < k0 + kb; ++c) {
                            const F lc = Lf[(size_t)c * n + j];
                            for (int i = c + lane; i < n; i += 64) Lf[(size_t)i * n + c] -= Lf[(size_t)i * n + j] * lc;
                        }
                        wave_sync();
                    }
                }
                __syncthreads();
                const int m0 = k0 + kb, nt = n - m0;
                for (int idx = tid; idx < nt * nt; idx += NT) {
                    const int r = m0 + idx / nt, c = m0 + idx % nt;
                    if (c <= r) {
                        F s = 0;
                        for (int mm = 0; mm < kb; ++mm) s += Lf[(size_t)r * n + k0 + mm] * Lf[(size_t)c * n + k0 + mm];
                        Lf[(size_t)r * n + c] -= s;
                    }
                }
                __syncthreads();
            }
            if (sflag[1] == 0) {
                // y = S^-1 b (wave 0), then FP64 refinement against S64
                for (int pass = 0; pass <= refine; ++pass) {
                    if (pass > 0) {
                        for (int i = tid; i < n; i += NT) {
                            double s = w.bc[i];
                            for (int k = 0; k < n; ++k) s -= w.S64[(size_t)i * n + k] * w.yc[k];
                            rsh[i] = s;
                        }
                        __syncthreads();
                    }
                    if (wave == 0) {
                        double yv[kWinNS];
#pragma unroll
                        for (int s = 0; s < kWinNS; ++s) {
                            const int i = 64 * s + lane;
                            yv[s] = i < n ? (pass == 0 ? w.bc[i] : rsh[i]) : 0.0;
                        }
                        wave_trsv<F, kWinNS>(Lf, n, yv, true);
                        wave_trsv<F, kWinNS>(Lf, n, yv, false);
#pragma unroll
                        for (int s = 0; s < kWinNS; ++s) {
                            const int i = 64 * s + lane;
                            if (i < n) w.yc[i] = pass == 0 ? yv[s] : w.yc[i] + yv[s];
                        }
                    }
                    __syncthreads();
                }
            }
        }
        const bool have_step = !invalid_blocks && sflag[1] == 0;
        __syncthreads();
        // ---------------- back-substitution and candidate ----------------
        double nrm[4] = {0.0, 0.0, 0.0, 0.0};  // |dx|^2, |x|^2, model, cost
        if (have_step) {
            for (int c = wave; c < K; c += NW) {
                double y[9];
                for (int i = 0; i < 9; ++i) y[i] = w.yc[9 * (size_t)c + i];
                for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                    const Real* rec = js + kWinJs * (size_t)j;
                    Real a0 = rec[33], a1 = rec[34];
#pragma unroll
                    for (int i = 0; i < 9; ++i) {
                        a0 -= rec[i] * (Real)y[i];
                        a1 -= rec[9 + i] * (Real)y[i];
                    }
                    R4<Real> uu;
                    uu.x = rec[18] * a0 + rec[21] * a1;
                    uu.y = rec[19] * a0 + rec[22] * a1;
                    uu.z = rec[20] * a0 + rec[23] * a1;
                    uu.w = (Real)0;
                    u[w.slot_of[j]] = uu;
                }
            }
            __syncthreads();
            for (int p = tid; p < P; p += NT) {
                const int ch = p >> 6, l = p & 63, len = w.chunk_len[ch];
                const size_t base = (size_t)w.chunk_off[ch];
                double t[3] = {0.0, 0.0, 0.0};
                for (int r = 0; r < len; ++r) {
                    const R4<Real> uu = u[base + 64 * (size_t)r + l];
                    t[0] += (double)uu.x;
                    t[1] += (double)uu.y;
                    t[2] += (double)uu.z;
                }
                double y[3];
                sym3_mul<double>(w.Vinv + 6 * (size_t)p, t, y);
                const double4 x4 = w.pt_x[p];
                const double xs[3] = {x4.x, x4.y, x4.z};
                double cand[3];
                for (int i = 0; i < 3; ++i) {
                    w.yp[3 * (size_t)p + i] = y[i];
                    const double dl = -w.sp[3 * (size_t)p + i] * y[i];
                    cand[i] = xs[i] + dl;
                    nrm[0] += dl * dl;
                    nrm[1] += xs[i] * xs[i];
                }
                w.pt_c[p] = make_double4(cand[0], cand[1], cand[2], 0.0);
            }
            for (int k = tid; k < n; k += NT) {
                const double dl = -w.sc[k] * w.yc[k];
                const double x = w.cam_x[k];
                w.cam_c[k] = x + dl;
                nrm[0] += dl * dl;
                nrm[1] += x * x;
            }
            __syncthreads();
            for (int c = tid; c < K; c += NT) {
                CameraPrep<double> Pc;
                camera_prep<double>(w.cam_c + 9 * (size_t)c, Pc);
                prep_c[c] = Pc;
            }
            __syncthreads();
            // model cost change from the stored J~ (step s = -y) and the FP64 candidate cost
            for (int c = wave; c < K; c += NW) {
                const CameraPrep<double> Pc = prep_c[c];
                double y[9];
                for (int i = 0; i < 9; ++i) y[i] = w.yc[9 * (size_t)c + i];
                for (int j = w.cam_off[c] + lane; j < w.cam_off[c + 1]; j += 64) {
                    const Real* rec = js + kWinJs * (size_t)j;
                    const int p = w.co_pt[j];
                    const double* yp = w.yp + 3 * (size_t)p;
                    double m0 = 0.0, m1 = 0.0;
                    for (int i = 0; i < 9; ++i) {
                        m0 -= (double)rec[i] * y[i];
                        m1 -= (double)rec[9 + i] * y[i];
                    }
                    for (int i = 0; i < 3; ++i) {
                        m0 -= (double)rec[18 + i] * yp[i];
                        m1 -= (double)rec[21 + i] * yp[i];
                    }
                    nrm[2] += m0 * ((double)rec[33] + 0.5 * m0) + m1 * ((double)rec[34] + 0.5 * m1);
                    const double4 x4 = w.pt_c[p];
                    const double X[3] = {x4.x, x4.y, x4.z};
                    double r[2];
                    snavely_residual<double>(Pc, X, w.co_uv64[2 * (size_t)j], w.co_uv64[2 * (size_t)j + 1], r);
                    nrm[3] += half_rho(r[0] * r[0] + r[1] * r[1], o.loss, o.loss_a);
                }
            }
        }
        wg_sum<4, NT>(nrm, red);
        // ---------------- decide ----------------
        if (tid == 0) {
            sflag[1] = 0;
            const double model = -red[2 * NW];
            double cand = red[3 * NW];
            const double dx2 = red[0], x2 = red[NW];
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
