// ba_oracle.cpp -- FP64 CPU restatement of the reference bundle-adjustment path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine: only tests/, bench.py's
// cpu_baseline leg and __graft_entry__.smoke() load it.  The product library never links, loads or
// calls it.  Built with -O2 -ffp-contract=off so every expression rounds exactly as written.
//
// What it restates (reference files are cited as path:line of luigifreda/slamplay):
//   * utils/geometry/rotation.h:57-67,133-198   dotProduct / crossProduct / angleAxisRotatePoint,
//                                                including the DBL_EPSILON first-order branch.
//   * utils/camera/SnavelyReprojectionError.h:42-87  residual = f*d*(-p0/p2, -p1/p2) - obs.
//   * ceres::AutoDiffCostFunction<...,2,9,3> (SnavelyReprojectionError.h:89-92): forward-mode
//     derivatives with a 12-lane dual number whose arithmetic follows Ceres's public Jet rules
//     (product rule f.a*g.v + f.v*g.a, quotient via 1/g.a, sqrt/sin/cos chain rules).
//   * ceres::HuberLoss(1.0) (test_bundle_adjustment_ceres.cpp:78) and Ceres's residual corrector
//     (rho'' <= 0 => r and J scaled by sqrt(rho')).
//   * ceres::Solve with Solver::Options defaults + SPARSE_SCHUR (test_bundle_adjustment_ceres.cpp:96-100):
//     trust-region Levenberg-Marquardt with Jacobi column scaling computed once at iteration 0,
//     D^2 = clamp(diag(J~'J~), 1e-6, 1e32)/radius, model cost change -(J~s).(r~ + J~s/2),
//     rho > 1e-3 acceptance, radius <- radius/max(1/3, 1-(2rho-1)^3), decrease factor doubling.
//     These Ceres semantics are restated from Ceres's public documentation; Ceres itself is not in
//     the reference tree (fetched by build_thirdparty.sh:190-197) so they are design inputs.
//   * The Schur complement reduced camera system solved by dense Cholesky (explicit S) or by
//     block-Jacobi preconditioned conjugate gradients on the implicit S.
//   * utils/datasets/BALProblem.cpp:39-76,230-336: RandNormal on glibc rand(), Median, Normalize,
//     Perturb, CameraToAngelAxisAndCenter, AngleAxisAndCenterToCamera.
//
// Pinned by tests/test_oracle.py against: 50-digit mpmath evaluations of the residual and its
// derivatives (tests/golden/snavely_golden.json), the Taylor-branch closed form d(RX)/dw = -[X]x,
// Huber closed forms, dense numpy solves of the full damped normal equations (Schur), PCG converging
// to the dense solution, noise-free convergence to zero cost, Normalize invariants and the
// rotation-only Perturb invariant.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

namespace oracle {
struct Options {
    int32_t max_num_iterations;
    int32_t max_num_consecutive_invalid_steps;
    int32_t jacobi_scaling;
    int32_t loss;          // 0 none, 1 Huber
    int32_t linear_solver; // 1 dense Schur, 2 iterative Schur (PCG)
    int32_t pcg_max_iterations;
    double loss_scale;
    double function_tolerance;
    double gradient_tolerance;
    double parameter_tolerance;
    double initial_trust_region_radius;
    double max_trust_region_radius;
    double min_trust_region_radius;
    double min_relative_decrease;
    double min_lm_diagonal;
    double max_lm_diagonal;
    double pcg_tolerance;
    int32_t mode;            // 0 Ceres driver, 1 g2o driver (test_bundle_adjustment_g2o.cpp)
    int32_t g2o_max_trials;  // g2o OptimizationAlgorithmLevenberg: rejected trials per iteration (10)
    double g2o_tau;          // g2o: lambda_0 = tau * max diag(H) (1e-5)
};

struct IterRecord {
    int32_t iteration;
    int32_t step_is_successful;
    int32_t step_is_valid;
    int32_t linear_solver_iterations;
    double cost;
    double cost_change;
    double gradient_max_norm;
    double step_norm;
    double relative_decrease;
    double trust_region_radius;
    double model_cost_change;
};

}  // namespace oracle

namespace {
using oracle::Options;
using oracle::IterRecord;

constexpr double kDblEps = std::numeric_limits<double>::epsilon();

// ---------------------------------------------------------------------------------------------
// 12-lane forward-mode dual number with Ceres's Jet arithmetic rules.
template <int N>
struct Jet {
    double a;
    double v[N];
    Jet() : a(0.0) { for (int i = 0; i < N; ++i) v[i] = 0.0; }
    explicit Jet(double x) : a(x) { for (int i = 0; i < N; ++i) v[i] = 0.0; }
    Jet(double x, int k) : a(x) { for (int i = 0; i < N; ++i) v[i] = 0.0; v[k] = 1.0; }
};

template <int N> inline Jet<N> operator+(const Jet<N>& f, const Jet<N>& g) {
    Jet<N> r(f.a + g.a); for (int i = 0; i < N; ++i) r.v[i] = f.v[i] + g.v[i]; return r;
}
template <int N> inline Jet<N> operator-(const Jet<N>& f, const Jet<N>& g) {
    Jet<N> r(f.a - g.a); for (int i = 0; i < N; ++i) r.v[i] = f.v[i] - g.v[i]; return r;
}
template <int N> inline Jet<N> operator-(const Jet<N>& f) {
    Jet<N> r(-f.a); for (int i = 0; i < N; ++i) r.v[i] = -f.v[i]; return r;
}
template <int N> inline Jet<N> operator*(const Jet<N>& f, const Jet<N>& g) {
    Jet<N> r(f.a * g.a); for (int i = 0; i < N; ++i) r.v[i] = f.a * g.v[i] + f.v[i] * g.a; return r;
}
template <int N> inline Jet<N> operator*(const Jet<N>& f, double s) {
    Jet<N> r(f.a * s); for (int i = 0; i < N; ++i) r.v[i] = f.v[i] * s; return r;
}
template <int N> inline Jet<N> operator/(const Jet<N>& f, const Jet<N>& g) {
    const double g_a_inverse = 1.0 / g.a;
    const double f_a_by_g_a = f.a * g_a_inverse;
    Jet<N> r(f_a_by_g_a);
    for (int i = 0; i < N; ++i) r.v[i] = (f.v[i] - f_a_by_g_a * g.v[i]) * g_a_inverse;
    return r;
}
template <int N> inline Jet<N> operator/(double s, const Jet<N>& g) {
    const double minus_s_g_a_inverse2 = -s / (g.a * g.a);
    Jet<N> r(s / g.a); for (int i = 0; i < N; ++i) r.v[i] = g.v[i] * minus_s_g_a_inverse2; return r;
}
template <int N> inline Jet<N> jsqrt(const Jet<N>& f) {
    const double tmp = std::sqrt(f.a);
    const double two_a_inverse = 1.0 / (2.0 * tmp);
    Jet<N> r(tmp); for (int i = 0; i < N; ++i) r.v[i] = f.v[i] * two_a_inverse; return r;
}
template <int N> inline Jet<N> jcos(const Jet<N>& f) {
    const double s = -std::sin(f.a);
    Jet<N> r(std::cos(f.a)); for (int i = 0; i < N; ++i) r.v[i] = s * f.v[i]; return r;
}
template <int N> inline Jet<N> jsin(const Jet<N>& f) {
    const double c = std::cos(f.a);
    Jet<N> r(std::sin(f.a)); for (int i = 0; i < N; ++i) r.v[i] = c * f.v[i]; return r;
}
template <int N> inline bool gt(const Jet<N>& f, double s) { return f.a > s; }
inline bool gt(double f, double s) { return f > s; }
inline double jsqrt(double x) { return std::sqrt(x); }
inline double jcos(double x) { return std::cos(x); }
inline double jsin(double x) { return std::sin(x); }

// ---------------------------------------------------------------------------------------------
// rotation.h:57-67
template <typename T> inline T dot3(const T x[3], const T y[3]) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; }
template <typename T> inline void cross3(const T x[3], const T y[3], T out[3]) {
    out[0] = x[1] * y[2] - x[2] * y[1];
    out[1] = x[2] * y[0] - x[0] * y[2];
    out[2] = x[0] * y[1] - x[1] * y[0];
}

// rotation.h:133-198 (Rodrigues; first-order branch when theta^2 <= DBL_EPSILON, whatever T is).
template <typename T>
void rotate_point(const T aa[3], const T pt[3], T out[3]) {
    const T theta2 = dot3(aa, aa);
    if (gt(theta2, kDblEps)) {
        const T theta = jsqrt(theta2);
        const T costheta = jcos(theta);
        const T sintheta = jsin(theta);
        const T theta_inverse = 1.0 / theta;
        const T w[3] = {aa[0] * theta_inverse, aa[1] * theta_inverse, aa[2] * theta_inverse};
        T w_cross_pt[3];
        cross3(w, pt, w_cross_pt);
        const T tmp = dot3(w, pt) * (T(1.0) - costheta);
        out[0] = pt[0] * costheta + w_cross_pt[0] * sintheta + w[0] * tmp;
        out[1] = pt[1] * costheta + w_cross_pt[1] * sintheta + w[1] * tmp;
        out[2] = pt[2] * costheta + w_cross_pt[2] * sintheta + w[2] * tmp;
    } else {
        T w_cross_pt[3];
        cross3(aa, pt, w_cross_pt);
        out[0] = pt[0] + w_cross_pt[0];
        out[1] = pt[1] + w_cross_pt[1];
        out[2] = pt[2] + w_cross_pt[2];
    }
}

// SnavelyReprojectionError.h:61-87
template <typename T>
void cam_projection(const T* camera, const T* point, T pred[2]) {
    T p[3];
    rotate_point(camera, point, p);
    p[0] = p[0] + camera[3];
    p[1] = p[1] + camera[4];
    p[2] = p[2] + camera[5];
    const T xp = -p[0] / p[2];
    const T yp = -p[1] / p[2];
    const T& l1 = camera[7];
    const T& l2 = camera[8];
    const T r2 = xp * xp + yp * yp;
    const T distortion = T(1.0) + r2 * (l1 + l2 * r2);
    const T& focal = camera[6];
    pred[0] = focal * distortion * xp;
    pred[1] = focal * distortion * yp;
}

// SnavelyReprojectionError.h:42-53
template <typename T>
void snavely_residual(const T* camera, const T* point, double ox, double oy, T r[2]) {
    T pred[2];
    cam_projection(camera, point, pred);
    r[0] = pred[0] - T(ox);
    r[1] = pred[1] - T(oy);
}

// Residual and 2x9 / 2x3 Jacobians (row-major per residual) of one observation by forward-mode
// differentiation, the AutoDiffCostFunction<...,2,9,3> semantics.
void residual_and_jacobian(const double* cam, const double* pt, double ox, double oy,
                           double r[2], double jc[18], double jp[6]) {
    typedef Jet<12> J;
    J c[9], x[3], res[2];
    for (int i = 0; i < 9; ++i) c[i] = J(cam[i], i);
    for (int i = 0; i < 3; ++i) x[i] = J(pt[i], 9 + i);
    snavely_residual(c, x, ox, oy, res);
    for (int k = 0; k < 2; ++k) {
        r[k] = res[k].a;
        if (jc) for (int i = 0; i < 9; ++i) jc[9 * k + i] = res[k].v[i];
        if (jp) for (int i = 0; i < 3; ++i) jp[3 * k + i] = res[k].v[9 + i];
    }
}

// g2o driver (test_bundle_adjustment_g2o.cpp:85): the rotation is updated as R <- exp(d) R, so its
// Jacobian columns are taken with respect to d: d p / d d_j = e_j x (R X).  The other columns are
// the autodiff ones (d p / d t = I is shared), so J_d = J_t [e_0 x q, e_1 x q, e_2 x q], q = R X.
void left_rotation_columns(const double* cam, const double* pt, double jc[18]) {
    double q[3];
    rotate_point(cam, pt, q);
    const double col[3][3] = {{0.0, -q[2], q[1]}, {q[2], 0.0, -q[0]}, {-q[1], q[0], 0.0}};  // col[j] = e_j x q
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 3; ++j)
            jc[9 * k + j] = jc[9 * k + 3] * col[j][0] + jc[9 * k + 4] * col[j][1] + jc[9 * k + 5] * col[j][2];
}

// SO(3) exp / log through unit quaternions (w, x, y, z) as Sophus defines them (the g2o driver's
// PoseAndIntrinsics, test_bundle_adjustment_g2o.cpp:48-63): exp(w) = (cos(t/2), sin(t/2) w/t) with
// series below t = 1e-10; log(q) = 2 atan(n/q_w) v/n with the small-n series and +-pi/n for q_w ~ 0.
void so3_exp(const double w[3], double q[4]) {
    const double t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double t = std::sqrt(t2);
    double re, im;
    if (t < 1e-10) {
        re = 1.0 - t2 / 8.0 + t2 * t2 / 384.0;
        im = 0.5 - t2 / 48.0 + t2 * t2 / 3840.0;
    } else {
        re = std::cos(0.5 * t);
        im = std::sin(0.5 * t) / t;
    }
    q[0] = re;
    q[1] = im * w[0];
    q[2] = im * w[1];
    q[3] = im * w[2];
}

void so3_log(const double q[4], double w[3]) {
    const double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const double n = std::sqrt(n2), qw = q[0];
    double k;
    if (n < 1e-10) k = 2.0 / qw - 2.0 * n2 / (3.0 * qw * qw * qw);
    else if (std::fabs(qw) < 1e-10) k = (qw >= 0.0 ? 3.14159265358979323846 : -3.14159265358979323846) / n;
    else k = 2.0 * std::atan(n / qw) / n;
    for (int i = 0; i < 3; ++i) w[i] = k * q[1 + i];
}

// angle-axis of exp(d) exp(w) (Hamilton product of the two quaternions)
void so3_left_update(const double d[3], const double w[3], double out[3]) {
    double a[4], b[4], c[4];
    so3_exp(d, a);
    so3_exp(w, b);
    c[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    c[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    c[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    c[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    so3_log(c, out);
}

// ceres::HuberLoss(a): rho(s) = s (s <= a^2), 2 a sqrt(s) - a^2 otherwise.
void huber(double s, double a, double rho[3]) {
    const double b = a * a;
    if (s > b) {
        const double r = std::sqrt(s);
        rho[0] = 2.0 * a * r - b;
        rho[1] = std::max(std::numeric_limits<double>::min(), a / r);
        rho[2] = -rho[1] / (2.0 * s);
    } else {
        rho[0] = s;
        rho[1] = 1.0;
        rho[2] = 0.0;
    }
}

// full_slam's stereo model (full_slam/include/myslam/g2o_types.h:122-175, backend.cpp:67-199): a
// pose T (angle-axis w, translation t in the 9-parameter slots, the last three unused) observes X
// through camera side s of the rig: p = E_s (R X + t), e = obs - (fx p_x/p_z + cx, fy p_y/p_z + cy).
struct StereoModel {
    double K[4];
    double R[2][9], t[2][3];
    const int* side;
};

struct PointObs;

// Timing aid for the CPU baseline: threads for the loops below that are split by camera / point so
// that every sum still runs over its observations in ascending order (results do not depend on the
// thread count).  1 unless oracle_set_threads says otherwise.
static int g_threads = 1;

struct Problem {
    int nc, np, no;
    const int* ci;
    const int* pi;
    const double* obs;
    const StereoModel* stereo = nullptr;
    // observations of each camera / point in ascending order; set for multi-threaded solves only
    const PointObs* by_cam = nullptr;
    const PointObs* by_pt = nullptr;
};

void stereo_point(const StereoModel& g, int s, const double* cam, const double* pt, double p[3]) {
    double q[3];
    rotate_point(cam, pt, q);
    for (int i = 0; i < 3; ++i) q[i] += cam[3 + i];
    for (int i = 0; i < 3; ++i) p[i] = g.R[s][3 * i] * q[0] + g.R[s][3 * i + 1] * q[1] + g.R[s][3 * i + 2] * q[2] + g.t[s][i];
}

void stereo_residual(const StereoModel& g, int s, const double* cam, const double* pt, double ox, double oy, double r[2]) {
    double p[3];
    stereo_point(g, s, cam, pt, p);
    r[0] = ox - (g.K[0] * p[0] / p[2] + g.K[2]);
    r[1] = oy - (g.K[1] * p[1] / p[2] + g.K[3]);
}

// The edge's Jacobian as written (g2o_types.h:147-172): pose columns (rho, phi) from p = E_s T X,
// Z^-1 taken as 1 / (Z + 1e-18); point columns J_rho E_s.R R.
void stereo_residual_jacobian(const StereoModel& g, int s, const double* cam, const double* pt, double ox, double oy,
                              double r[2], double jc[18], double jp[6]) {
    double p[3];
    stereo_point(g, s, cam, pt, p);
    r[0] = ox - (g.K[0] * p[0] / p[2] + g.K[2]);
    r[1] = oy - (g.K[1] * p[1] / p[2] + g.K[3]);
    const double fx = g.K[0], fy = g.K[1], X = p[0], Y = p[1], Z = p[2];
    const double zi = 1.0 / (Z + 1e-18), zi2 = zi * zi;
    const double row0[6] = {-fx * zi, 0.0, fx * X * zi2, fx * X * Y * zi2, -fx - fx * X * X * zi2, fx * Y * zi};
    const double row1[6] = {0.0, -fy * zi, fy * Y * zi2, fy + fy * Y * Y * zi2, -fy * X * Y * zi2, -fy * X * zi};
    for (int i = 0; i < 9; ++i) {
        jc[i] = i < 6 ? row0[i] : 0.0;
        jc[9 + i] = i < 6 ? row1[i] : 0.0;
    }
    double Rt[9];  // R of the pose: columns R e_k
    for (int k = 0; k < 3; ++k) {
        double e[3] = {0.0, 0.0, 0.0}, c[3];
        e[k] = 1.0;
        rotate_point(cam, e, c);
        for (int i = 0; i < 3; ++i) Rt[3 * i + k] = c[i];
    }
    for (int row = 0; row < 2; ++row)
        for (int k = 0; k < 3; ++k) {
            double v = 0.0;
            for (int i = 0; i < 3; ++i) {
                double er = 0.0;  // (E_s.R R)[i][k]
                for (int m = 0; m < 3; ++m) er += g.R[s][3 * i + m] * Rt[3 * m + k];
                v += jc[9 * row + i] * er;
            }
            jp[3 * row + k] = v;
        }
}

// T <- exp(xi) T, xi = (rho, phi): R' = exp(phi) R, t' = exp(phi) t + V(phi) rho with the SO(3) left
// Jacobian V = I + (1 - cos a)/a^2 [phi]x + (a - sin a)/a^3 [phi]x^2 (Sophus SE3::exp, which
// VertexPose::oplusImpl applies, full_slam/include/myslam/g2o_types.h:55-60).
void se3_left_update(const double* xi, const double* w, const double* t, double* w_out, double* t_out) {
    const double* rho = xi;
    const double* phi = xi + 3;
    so3_left_update(phi, w, w_out);
    double Rt[3];
    rotate_point(phi, t, Rt);
    const double a2 = phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2], a = std::sqrt(a2);
    // (1 - cos a)/a^2 as 2 sin^2(a/2)/a^2 and (a - sin a)/a^3 by its Taylor series below 0.01: the
    // direct quotients lose up to half the digits to cancellation at small angles
    const double sh = std::sin(0.5 * a);
    const double k1 = a < 1e-150 ? 0.5 : 2.0 * sh * sh / a2;
    const double k2 = a < 1e-2 ? 1.0 / 6.0 - a2 / 120.0 + a2 * a2 / 5040.0 : (a - std::sin(a)) / (a2 * a);
    double pr[3], ppr[3];
    cross3(phi, rho, pr);
    cross3(phi, pr, ppr);
    for (int i = 0; i < 3; ++i) t_out[i] = Rt[i] + (rho[i] + k1 * pr[i] + k2 * ppr[i]);
}

// Observations' lists (CSR) per camera or per point, ascending within each list.
struct PointObs {
    std::vector<int> off, idx;
};

PointObs obs_lists(const int* key, int n_keys, int no) {
    PointObs po;
    po.off.assign(n_keys + 1, 0);
    for (int k = 0; k < no; ++k) po.off[key[k] + 1]++;
    for (int c = 0; c < n_keys; ++c) po.off[c + 1] += po.off[c];
    po.idx.resize(no);
    std::vector<int> fill(po.off.begin(), po.off.end() - 1);
    for (int k = 0; k < no; ++k) po.idx[fill[key[k]]++] = k;
    return po;
}

// f(k) for every observation, grouped by camera (by_camera) or by point: each camera's / point's
// observations in ascending order, the groups split over threads when the lists are present.  A
// sum that f accumulates into its camera's / point's slot therefore runs in the serial order.
template <typename F>
void for_obs_by(const Problem& P, bool by_camera, F&& f) {
    const PointObs* lists = by_camera ? P.by_cam : P.by_pt;
    if (!lists || g_threads <= 1) {
        for (int k = 0; k < P.no; ++k) f(k);
        return;
    }
    const int n = by_camera ? P.nc : P.np;
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads)
    for (int g = 0; g < n; ++g)
        for (int a = lists->off[g]; a < lists->off[g + 1]; ++a) f(lists->idx[a]);
}

// Per-observation linearization (corrected, unscaled) plus problem-level gradient and cost.
struct Linearization {
    std::vector<double> jc, jp, r;  // 18, 6, 2 per observation (Huber-corrected)
    std::vector<double> grad;       // 9nc + 3np, J'r (unscaled)
    double cost = 0.0;
};

double obs_cost(const Options& o, double r0, double r1, double* sqrt_rho1) {
    const double s = r0 * r0 + r1 * r1;
    if (o.loss == 1) {
        double rho[3];
        huber(s, o.loss_scale, rho);
        if (sqrt_rho1) *sqrt_rho1 = std::sqrt(rho[1]);
        return 0.5 * rho[0];
    }
    if (sqrt_rho1) *sqrt_rho1 = 1.0;
    return 0.5 * s;
}

// Cost only (Ceres Evaluate with no Jacobian). Returns false on a non-finite cost.
bool evaluate_cost(const Problem& P, const Options& o, const double* cams, const double* pts, double* cost) {
    std::vector<double> ck(P.no);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int k = 0; k < P.no; ++k) {
        double r[2];
        if (P.stereo)
            stereo_residual(*P.stereo, P.stereo->side[k], cams + 9 * P.ci[k], pts + 3 * P.pi[k], P.obs[2 * k], P.obs[2 * k + 1], r);
        else
            snavely_residual(cams + 9 * P.ci[k], pts + 3 * P.pi[k], P.obs[2 * k], P.obs[2 * k + 1], r);
        ck[k] = obs_cost(o, r[0], r[1], nullptr);
    }
    double sum = 0.0;
    for (int k = 0; k < P.no; ++k) sum += ck[k];
    *cost = sum;
    return std::isfinite(sum);
}

bool linearize(const Problem& P, const Options& o, const double* cams, const double* pts, Linearization* L) {
    L->jc.assign(18 * (size_t)P.no, 0.0);
    L->jp.assign(6 * (size_t)P.no, 0.0);
    L->r.assign(2 * (size_t)P.no, 0.0);
    L->grad.assign(9 * (size_t)P.nc + 3 * (size_t)P.np, 0.0);
    std::vector<double> ck(P.no);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int k = 0; k < P.no; ++k) {
        double r[2], jc[18], jp[6], w;
        if (P.stereo) {
            stereo_residual_jacobian(*P.stereo, P.stereo->side[k], cams + 9 * P.ci[k], pts + 3 * P.pi[k], P.obs[2 * k],
                                     P.obs[2 * k + 1], r, jc, jp);
        } else {
            residual_and_jacobian(cams + 9 * P.ci[k], pts + 3 * P.pi[k], P.obs[2 * k], P.obs[2 * k + 1], r, jc, jp);
            if (o.mode == 1) left_rotation_columns(cams + 9 * P.ci[k], pts + 3 * P.pi[k], jc);
        }
        ck[k] = obs_cost(o, r[0], r[1], &w);
        for (int i = 0; i < 18; ++i) L->jc[18 * (size_t)k + i] = w * jc[i];
        for (int i = 0; i < 6; ++i) L->jp[6 * (size_t)k + i] = w * jp[i];
        L->r[2 * (size_t)k] = w * r[0];
        L->r[2 * (size_t)k + 1] = w * r[1];
    }
    double sum = 0.0;
    for (int k = 0; k < P.no; ++k) sum += ck[k];
    auto grad_cam = [&](int k) {
        double* gc = &L->grad[9 * (size_t)P.ci[k]];
        const double* J = &L->jc[18 * (size_t)k];
        const double* rr = &L->r[2 * (size_t)k];
        for (int i = 0; i < 9; ++i) gc[i] += J[i] * rr[0] + J[9 + i] * rr[1];
    };
    auto grad_pt = [&](int k) {
        double* gp = &L->grad[9 * (size_t)P.nc + 3 * (size_t)P.pi[k]];
        const double* Jp = &L->jp[6 * (size_t)k];
        const double* rr = &L->r[2 * (size_t)k];
        for (int i = 0; i < 3; ++i) gp[i] += Jp[i] * rr[0] + Jp[3 + i] * rr[1];
    };
    for_obs_by(P, true, grad_cam);
    for_obs_by(P, false, grad_pt);
    L->cost = sum;
    return std::isfinite(sum);
}

// In-place lower Cholesky of an n x n row-major SPD matrix; false if a pivot is not positive.
bool cholesky(double* A, int n) {
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        d = std::sqrt(d);
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s / d;
        }
    }
    return true;
}

void cholesky_solve(const double* L, int n, double* b) {
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * b[k];
        b[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * b[k];
        b[i] = s / L[(size_t)i * n + i];
    }
}

// Inverse of an SPD matrix through its Cholesky factor; false if not positive definite.
bool spd_inverse(const double* A, int n, double* Ainv) {
    std::vector<double> L(A, A + (size_t)n * n);
    if (!cholesky(L.data(), n)) return false;
    for (int j = 0; j < n; ++j) {
        std::vector<double> e(n, 0.0);
        e[j] = 1.0;
        cholesky_solve(L.data(), n, e.data());
        for (int i = 0; i < n; ++i) Ainv[(size_t)i * n + j] = e[i];
    }
    return true;
}

// Reduced camera system pieces shared by the dense and iterative Schur solvers. All quantities are
// in the Jacobi-scaled variables.
struct Schur {
    int nc, np;
    std::vector<double> U;     // 81 per camera: J~c'J~c + Dc^2
    std::vector<double> Vinv;  // 9 per point: (J~p'J~p + Dp^2)^-1
    std::vector<double> W;     // 27 per observation: J~c'J~p (9x3 row-major)
    std::vector<double> gc, gp;
};

// Build damped blocks. Returns false if a point block is not positive definite.
bool build_blocks(const Problem& P, const Linearization& L, const std::vector<double>& scale,
                  const std::vector<double>& diag_d2, Schur* S) {
    const int nc = P.nc, np = P.np;
    S->nc = nc;
    S->np = np;
    S->U.assign(81 * (size_t)nc, 0.0);
    std::vector<double> V(9 * (size_t)np, 0.0);
    S->W.assign(27 * (size_t)P.no, 0.0);
    S->gc.assign(9 * (size_t)nc, 0.0);
    S->gp.assign(3 * (size_t)np, 0.0);
    auto scaled = [&](int k, double jc[18], double jp[6]) {
        const int c = P.ci[k], p = P.pi[k];
        for (int row = 0; row < 2; ++row) {
            for (int i = 0; i < 9; ++i) jc[9 * row + i] = L.jc[18 * (size_t)k + 9 * row + i] * scale[9 * (size_t)c + i];
            for (int i = 0; i < 3; ++i) jp[3 * row + i] = L.jp[6 * (size_t)k + 3 * row + i] * scale[9 * (size_t)nc + 3 * (size_t)p + i];
        }
    };
    auto cam_part = [&](int k) {  // U and g_c of the observation's camera
        double jc[18], jp[6];
        scaled(k, jc, jp);
        const int c = P.ci[k];
        const double* r = &L.r[2 * (size_t)k];
        double* U = &S->U[81 * (size_t)c];
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 9; ++j) U[9 * i + j] += jc[i] * jc[j] + jc[9 + i] * jc[9 + j];
        for (int i = 0; i < 9; ++i) S->gc[9 * (size_t)c + i] += jc[i] * r[0] + jc[9 + i] * r[1];
    };
    auto pt_part = [&](int k) {  // V, W and g_p of the observation's point
        double jc[18], jp[6];
        scaled(k, jc, jp);
        const int p = P.pi[k];
        const double* r = &L.r[2 * (size_t)k];
        double* Vp = &V[9 * (size_t)p];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Vp[3 * i + j] += jp[i] * jp[j] + jp[3 + i] * jp[3 + j];
        double* W = &S->W[27 * (size_t)k];
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 3; ++j) W[3 * i + j] = jc[i] * jp[j] + jc[9 + i] * jp[3 + j];
        for (int i = 0; i < 3; ++i) S->gp[3 * (size_t)p + i] += jp[i] * r[0] + jp[3 + i] * r[1];
    };
    for_obs_by(P, true, cam_part);
    for_obs_by(P, false, pt_part);
    for (int c = 0; c < nc; ++c)
        for (int i = 0; i < 9; ++i) S->U[81 * (size_t)c + 10 * i] += diag_d2[9 * (size_t)c + i];
    S->Vinv.assign(9 * (size_t)np, 0.0);
    for (int p = 0; p < np; ++p) {
        double* Vp = &V[9 * (size_t)p];
        for (int i = 0; i < 3; ++i) Vp[4 * i] += diag_d2[9 * (size_t)nc + 3 * (size_t)p + i];
        if (!spd_inverse(Vp, 3, &S->Vinv[9 * (size_t)p])) return false;
    }
    return true;
}

// Diagonal of J~'J~ (scaled, corrected Jacobian).
void jtj_diagonal(const Problem& P, const Linearization& L, const std::vector<double>& scale, std::vector<double>* d) {
    d->assign(9 * (size_t)P.nc + 3 * (size_t)P.np, 0.0);
    for_obs_by(P, true, [&](int k) {
        const int c = P.ci[k];
        for (int i = 0; i < 9; ++i) {
            const double s = scale[9 * (size_t)c + i];
            const double a = L.jc[18 * (size_t)k + i] * s, b = L.jc[18 * (size_t)k + 9 + i] * s;
            (*d)[9 * (size_t)c + i] += a * a + b * b;
        }
    });
    for_obs_by(P, false, [&](int k) {
        const int p = P.pi[k];
        for (int i = 0; i < 3; ++i) {
            const size_t col = 9 * (size_t)P.nc + 3 * (size_t)p + i;
            const double s = scale[col];
            const double a = L.jp[6 * (size_t)k + i] * s, b = L.jp[6 * (size_t)k + 3 + i] * s;
            (*d)[col] += a * a + b * b;
        }
    });
}

// Points' observation lists (CSR) for the Schur elimination.
PointObs point_obs(const Problem& P) { return obs_lists(P.pi, P.np, P.no); }

// q = Vinv_p * v (3x3 times 3)
inline void mul33(const double* M, const double* v, double* out) {
    for (int i = 0; i < 3; ++i) out[i] = M[3 * i] * v[0] + M[3 * i + 1] * v[1] + M[3 * i + 2] * v[2];
}

// rhs_c = gc - sum_k W_k Vinv_p gp
void schur_rhs(const Problem& P, const Schur& S, std::vector<double>* b) {
    b->assign(S.gc.begin(), S.gc.end());
    std::vector<double> q(3 * (size_t)P.np);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int p = 0; p < P.np; ++p) mul33(&S.Vinv[9 * (size_t)p], &S.gp[3 * (size_t)p], &q[3 * (size_t)p]);
    for_obs_by(P, true, [&](int k) {
        const double* W = &S.W[27 * (size_t)k];
        const double* qp = &q[3 * (size_t)P.pi[k]];
        double* bc = &(*b)[9 * (size_t)P.ci[k]];
        for (int i = 0; i < 9; ++i) bc[i] -= W[3 * i] * qp[0] + W[3 * i + 1] * qp[1] + W[3 * i + 2] * qp[2];
    });
}

// yp = Vinv_p (gp - sum_k W_k' yc)
void schur_back_substitute(const Problem& P, const Schur& S, const double* yc, double* yp) {
    std::vector<double> t(S.gp);
    for_obs_by(P, false, [&](int k) {
        const double* W = &S.W[27 * (size_t)k];
        const double* y = yc + 9 * (size_t)P.ci[k];
        double* tp = &t[3 * (size_t)P.pi[k]];
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int i = 0; i < 9; ++i) s += W[3 * i + j] * y[i];
            tp[j] -= s;
        }
    });
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int p = 0; p < P.np; ++p) mul33(&S.Vinv[9 * (size_t)p], &t[3 * (size_t)p], yp + 3 * (size_t)p);
}

// Dense explicit reduced camera matrix S = U* - sum_p W Vinv W'.
void schur_dense_matrix(const Problem& P, const PointObs& po, const Schur& S, std::vector<double>* A) {
    const int n = 9 * P.nc;
    A->assign((size_t)n * n, 0.0);
    for (int c = 0; c < P.nc; ++c)
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 9; ++j) (*A)[(size_t)(9 * c + i) * n + 9 * c + j] = S.U[81 * (size_t)c + 9 * i + j];
    std::vector<double> Y(27);
    for (int p = 0; p < P.np; ++p) {
        const double* Vi = &S.Vinv[9 * (size_t)p];
        for (int a = po.off[p]; a < po.off[p + 1]; ++a) {
            const int ka = po.idx[a];
            const double* Wa = &S.W[27 * (size_t)ka];
            for (int i = 0; i < 9; ++i)
                for (int j = 0; j < 3; ++j)
                    Y[3 * i + j] = Wa[3 * i] * Vi[j] + Wa[3 * i + 1] * Vi[3 + j] + Wa[3 * i + 2] * Vi[6 + j];
            const int ca = P.ci[ka];
            for (int b = po.off[p]; b < po.off[p + 1]; ++b) {
                const int kb = po.idx[b];
                const double* Wb = &S.W[27 * (size_t)kb];
                const int cb = P.ci[kb];
                for (int i = 0; i < 9; ++i)
                    for (int j = 0; j < 9; ++j)
                        (*A)[(size_t)(9 * ca + i) * n + 9 * cb + j] -=
                            Y[3 * i] * Wb[3 * j] + Y[3 * i + 1] * Wb[3 * j + 1] + Y[3 * i + 2] * Wb[3 * j + 2];
            }
        }
    }
}

// out = S v for the implicit reduced camera matrix.
void schur_matvec(const Problem& P, const Schur& S, const double* v, double* out) {
    std::vector<double> t(3 * (size_t)P.np, 0.0), z(3 * (size_t)P.np);
    for_obs_by(P, false, [&](int k) {
        const double* W = &S.W[27 * (size_t)k];
        const double* vc = v + 9 * (size_t)P.ci[k];
        double* tp = &t[3 * (size_t)P.pi[k]];
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int i = 0; i < 9; ++i) s += W[3 * i + j] * vc[i];
            tp[j] += s;
        }
    });
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int p = 0; p < P.np; ++p) mul33(&S.Vinv[9 * (size_t)p], &t[3 * (size_t)p], &z[3 * (size_t)p]);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int c = 0; c < P.nc; ++c) {
        const double* U = &S.U[81 * (size_t)c];
        for (int i = 0; i < 9; ++i) {
            double s = 0.0;
            for (int j = 0; j < 9; ++j) s += U[9 * i + j] * v[9 * (size_t)c + j];
            out[9 * (size_t)c + i] = s;
        }
    }
    for_obs_by(P, true, [&](int k) {
        const double* W = &S.W[27 * (size_t)k];
        const double* zp = &z[3 * (size_t)P.pi[k]];
        double* oc = out + 9 * (size_t)P.ci[k];
        for (int i = 0; i < 9; ++i) oc[i] -= W[3 * i] * zp[0] + W[3 * i + 1] * zp[1] + W[3 * i + 2] * zp[2];
    });
}

double dotv(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// Block-Jacobi preconditioned CG on the implicit S, x0 = 0, stop when ||r|| <= tol ||b||.
// Returns the iteration count, or -1 if a preconditioner block is not positive definite.
int schur_pcg(const Problem& P, const Schur& S, const std::vector<double>& b, int max_iters, double tol,
              std::vector<double>* x) {
    const int nc = P.nc;
    const size_t n = 9 * (size_t)nc;
    // M_c = U*_c - sum_{k in c} W_k Vinv_p W_k'
    std::vector<double> M(S.U), Minv(81 * (size_t)nc);
    for_obs_by(P, true, [&](int k) {
        const double* W = &S.W[27 * (size_t)k];
        const double* Vi = &S.Vinv[9 * (size_t)P.pi[k]];
        double Y[27];
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 3; ++j) Y[3 * i + j] = W[3 * i] * Vi[j] + W[3 * i + 1] * Vi[3 + j] + W[3 * i + 2] * Vi[6 + j];
        double* Mc = &M[81 * (size_t)P.ci[k]];
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 9; ++j) Mc[9 * i + j] -= Y[3 * i] * W[3 * j] + Y[3 * i + 1] * W[3 * j + 1] + Y[3 * i + 2] * W[3 * j + 2];
    });
    bool spd = true;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1) reduction(&& : spd)
    for (int c = 0; c < nc; ++c) spd = spd_inverse(&M[81 * (size_t)c], 9, &Minv[81 * (size_t)c]) && spd;
    if (!spd) return -1;
    auto precond = [&](const std::vector<double>& r, std::vector<double>* z) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (int c = 0; c < nc; ++c) {
            const double* Mi = &Minv[81 * (size_t)c];
            for (int i = 0; i < 9; ++i) {
                double s = 0.0;
                for (int j = 0; j < 9; ++j) s += Mi[9 * i + j] * r[9 * (size_t)c + j];
                (*z)[9 * (size_t)c + i] = s;
            }
        }
    };
    x->assign(n, 0.0);
    std::vector<double> r(b), z(n), p(n), q(n);
    const double bnorm = std::sqrt(dotv(b, b));
    if (bnorm == 0.0) return 0;
    precond(r, &z);
    p = z;
    double rz = dotv(r, z);
    int it = 0;
    while (it < max_iters) {
        schur_matvec(P, S, p.data(), q.data());
        const double alpha = rz / dotv(p, q);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (size_t i = 0; i < n; ++i) {
            (*x)[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        ++it;
        if (std::sqrt(dotv(r, r)) <= tol * bnorm) break;
        precond(r, &z);
        const double rz_new = dotv(r, z);
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return it;
}

// One LM linear solve at the current linearization.  Output: scaled y with (J~'J~ + D^2) y = J~'r~.
// Returns linear-solver iterations (1 for dense), or -1 on failure.
int solve_damped(const Problem& P, const Options& o, const PointObs& po, const Linearization& L,
                 const std::vector<double>& scale, const std::vector<double>& diag, double radius,
                 std::vector<double>* y) {
    const size_t nparam = 9 * (size_t)P.nc + 3 * (size_t)P.np;
    std::vector<double> d2(nparam);
    for (size_t i = 0; i < nparam; ++i) d2[i] = diag[i] / radius;
    Schur S;
    if (!build_blocks(P, L, scale, d2, &S)) return -1;
    std::vector<double> b;
    schur_rhs(P, S, &b);
    std::vector<double> yc;
    int iters = 1;
    if (o.linear_solver == 2) {
        iters = schur_pcg(P, S, b, o.pcg_max_iterations, o.pcg_tolerance, &yc);
        if (iters < 0) return -1;
    } else {
        std::vector<double> A;
        schur_dense_matrix(P, po, S, &A);
        if (!cholesky(A.data(), 9 * P.nc)) return -1;
        yc = b;
        cholesky_solve(A.data(), 9 * P.nc, yc.data());
    }
    y->assign(nparam, 0.0);
    std::copy(yc.begin(), yc.end(), y->begin());
    schur_back_substitute(P, S, yc.data(), y->data() + 9 * (size_t)P.nc);
    for (size_t i = 0; i < nparam; ++i)
        if (!std::isfinite((*y)[i])) return -1;
    return iters;
}

// model_cost_change = -(J~ s).(r~ + J~ s / 2), s the scaled step.
double model_cost_change(const Problem& P, const Linearization& L, const std::vector<double>& delta) {
    std::vector<double> term(2 * (size_t)P.no);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int k = 0; k < P.no; ++k) {
        const double* jc = &L.jc[18 * (size_t)k];
        const double* jp = &L.jp[6 * (size_t)k];
        const double* dc = &delta[9 * (size_t)P.ci[k]];
        const double* dp = &delta[9 * (size_t)P.nc + 3 * (size_t)P.pi[k]];
        for (int row = 0; row < 2; ++row) {
            double m = 0.0;
            for (int i = 0; i < 9; ++i) m += jc[9 * row + i] * dc[i];
            for (int i = 0; i < 3; ++i) m += jp[3 * row + i] * dp[i];
            term[2 * (size_t)k + row] = m * (L.r[2 * (size_t)k + row] + 0.5 * m);
        }
    }
    double sum = 0.0;
    for (size_t i = 0; i < term.size(); ++i) sum += term[i];
    return -sum;
}

double norm2(const double* a, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * a[i];
    return std::sqrt(s);
}

// ---------------------------------------------------------------------------------------------
// BALProblem.cpp:39-55 (glibc rand(), Marsaglia polar, returns x1*w)
double rand_double() { return static_cast<double>(rand()) / RAND_MAX; }
double rand_normal() {
    double x1, x2, w;
    do {
        x1 = 2.0 * rand_double() - 1.0;
        x2 = 2.0 * rand_double() - 1.0;
        w = x1 * x1 + x2 * x2;
    } while (w >= 1.0 || w == 0.0);
    w = std::sqrt((-2.0 * std::log(w)) / w);
    return x1 * w;
}
// BALProblem.cpp:71-76 (upper median via nth_element)
double median(std::vector<double>* d) {
    const int n = (int)d->size();
    auto mid = d->begin() + n / 2;
    std::nth_element(d->begin(), mid, d->end());
    return *mid;
}
// BALProblem.cpp:230-246
void camera_to_aa_center(const double* cam, double aa[3], double center[3]) {
    for (int i = 0; i < 3; ++i) aa[i] = cam[i];
    const double inv[3] = {-aa[0], -aa[1], -aa[2]};
    rotate_point(inv, cam + 3, center);
    for (int i = 0; i < 3; ++i) center[i] *= -1.0;
}
// BALProblem.cpp:248-261
void aa_center_to_camera(const double aa[3], const double center[3], double* cam) {
    for (int i = 0; i < 3; ++i) cam[i] = aa[i];
    rotate_point(aa, center, cam + 3);
    for (int i = 0; i < 3; ++i) cam[3 + i] *= -1.0;
}

}  // namespace

// =============================================================================================
extern "C" {

int oracle_abi_version(void) { return 1; }

// Raw (uncorrected) residuals and Jacobians, original observation order; Jc 2x9 and Jp 2x3 row-major.
int oracle_eval(int no, const int* ci, const int* pi, const double* obs, const double* cams, const double* pts,
                double* residuals, double* jc, double* jp) {
    for (int k = 0; k < no; ++k)
        residual_and_jacobian(cams + 9 * ci[k], pts + 3 * pi[k], obs[2 * k], obs[2 * k + 1], residuals + 2 * (size_t)k,
                              jc ? jc + 18 * (size_t)k : nullptr, jp ? jp + 6 * (size_t)k : nullptr);
    return 0;
}

// Residual of one observation in plain double arithmetic (no dual numbers).
void oracle_residual(const double* cam, const double* pt, double ox, double oy, double r[2]) {
    snavely_residual(cam, pt, ox, oy, r);
}

void oracle_rotate_point(const double* aa, const double* pt, double* out) { rotate_point(aa, pt, out); }

// g2o-mode Jacobian of one observation: rotation columns with respect to the left update exp(d) R.
void oracle_jacobian_left(const double* cam, const double* pt, double ox, double oy, double* r, double* jc, double* jp) {
    residual_and_jacobian(cam, pt, ox, oy, r, jc, jp);
    left_rotation_columns(cam, pt, jc);
}

// angle-axis of exp(d) exp(w)
void oracle_so3_left_update(const double* d, const double* w, double* out) { so3_left_update(d, w, out); }

// Threads for oracle_solve (a timing aid; results do not depend on it).  Returns the count in use.
int oracle_set_threads(int n) {
#ifdef _OPENMP
    g_threads = n < 1 ? 1 : n;
#else
    (void)n;
    g_threads = 1;
#endif
    return g_threads;
}

void oracle_se3_left_update(const double* xi, const double* w, const double* t, double* w_out, double* t_out) {
    se3_left_update(xi, w, t, w_out, t_out);
}

// rig from K and the two extrinsics (angle-axis, translation)
static StereoModel make_rig(const double* K, const double* ext, const int* side) {
    StereoModel g{};
    for (int i = 0; i < 4; ++i) g.K[i] = K[i];
    for (int s = 0; s < 2; ++s) {
        for (int k = 0; k < 3; ++k) {
            double e[3] = {0.0, 0.0, 0.0}, c[3];
            e[k] = 1.0;
            rotate_point(ext + 6 * s, e, c);
            for (int i = 0; i < 3; ++i) g.R[s][3 * i + k] = c[i];
        }
        for (int i = 0; i < 3; ++i) g.t[s][i] = ext[6 * s + 3 + i];
    }
    g.side = side;
    return g;
}

void oracle_stereo_residual_jacobian(const double* cam, const double* pt, const double* K, const double* ext, int side,
                                     double ox, double oy, double* r, double* jc, double* jp) {
    const StereoModel g = make_rig(K, ext, nullptr);
    stereo_residual_jacobian(g, side, cam, pt, ox, oy, r, jc, jp);
}

void oracle_stereo_residual(const double* cam, const double* pt, const double* K, const double* ext, int side, double ox,
                            double oy, double* r) {
    const StereoModel g = make_rig(K, ext, nullptr);
    stereo_residual(g, side, cam, pt, ox, oy, r);
}

void oracle_huber(double s, double a, double* rho) { huber(s, a, rho); }

// Total cost 1/2 sum rho(|r|^2) in observation order.
double oracle_cost(int no, const int* ci, const int* pi, const double* obs, const double* cams, const double* pts,
                   int loss, double loss_scale) {
    Options o{};
    o.loss = loss;
    o.loss_scale = loss_scale;
    Problem P{0, 0, no, ci, pi, obs};
    double c;
    evaluate_cost(P, o, cams, pts, &c);
    return c;
}

// One damped LM step at x for a given radius, with Jacobi scaling computed at x (iteration-0 rule).
// Outputs the unscaled parameter update delta (9nc + 3np) and the model cost change; returns the
// linear-solver iteration count or -1.
int oracle_lm_step(int nc, int np, int no, const int* ci, const int* pi, const double* obs, const double* cams,
                   const double* pts, const Options* o, double radius, double* delta_out, double* model_change,
                   double* scale_out) {
    Problem P{nc, np, no, ci, pi, obs};
    Linearization L;
    linearize(P, *o, cams, pts, &L);
    const size_t nparam = 9 * (size_t)nc + 3 * (size_t)np;
    std::vector<double> scale(nparam, 1.0), ones(nparam, 1.0), diag;
    if (o->jacobi_scaling) {
        jtj_diagonal(P, L, ones, &diag);
        for (size_t i = 0; i < nparam; ++i) scale[i] = 1.0 / (1.0 + std::sqrt(diag[i]));
    }
    jtj_diagonal(P, L, scale, &diag);
    for (size_t i = 0; i < nparam; ++i) diag[i] = std::min(std::max(diag[i], o->min_lm_diagonal), o->max_lm_diagonal);
    PointObs po = point_obs(P);
    std::vector<double> y;
    const int it = solve_damped(P, *o, po, L, scale, diag, radius, &y);
    if (it < 0) return -1;
    std::vector<double> delta(nparam);
    for (size_t i = 0; i < nparam; ++i) delta[i] = -y[i] * scale[i];
    if (delta_out) std::copy(delta.begin(), delta.end(), delta_out);
    if (scale_out) std::copy(scale.begin(), scale.end(), scale_out);
    if (model_change) *model_change = model_cost_change(P, L, delta);
    return it;
}

// Ceres-style trust-region LM.  cams/pts are updated in place with the best parameters found.
// termination: 0 CONVERGENCE, 1 NO_CONVERGENCE, 2 FAILURE.
// The g2o driver's solve (test_bundle_adjustment_g2o.cpp:166-218: OptimizationAlgorithmLevenberg,
// optimize(40), marginalized points, CSparse) on the same residual.  g2o 20230223 is not in the
// reference tree (build_thirdparty.sh:253); its Levenberg algorithm is restated from its public
// source: lambda_0 = tau * max_i H_ii; per iteration up to max_trials trials of
// (H + lambda I) x = b, x applied through the vertices' oplus (rotation exp(x) R, the rest additive),
// rho = (chi2 - chi2_new) / (x'(lambda x + b) + 1e-3); rho > 0: accept, lambda *= max(1/3,
// min(2/3, 1 - (2 rho - 1)^3)), nu = 2; else lambda *= nu, nu *= 2; the solve stops after max_trials
// rejections (or rho == 0).  chi2 = sum rho(s) = 2 cost; here x'(lambda x + b) = 2 * model cost
// change, which is exact for the solved damped system.  Radius = 1 / lambda in the trace; costs in
// the records are chi2.
static int solve_g2o(const Problem& P, const Options& o0, double* cams, double* pts, IterRecord* trace, int trace_cap,
                     int* n_trace, double* initial_cost, double* final_cost, int* termination) {
    Options o = o0;
    o.jacobi_scaling = 0;
    o.min_lm_diagonal = o.max_lm_diagonal = 1.0;
    const size_t ncam = 9 * (size_t)P.nc, nparam = ncam + 3 * (size_t)P.np;
    std::vector<double> x(nparam), cand(nparam);
    std::copy(cams, cams + ncam, x.begin());
    std::copy(pts, pts + 3 * (size_t)P.np, x.begin() + ncam);
    int nt = 0;
    auto push = [&](const IterRecord& r) { if (trace && nt < trace_cap) trace[nt] = r; ++nt; };
    PointObs po = point_obs(P);
    Linearization L;
    if (!linearize(P, o, x.data(), x.data() + ncam, &L)) {
        *termination = 2;
        if (n_trace) *n_trace = 0;
        return 2;
    }
    const std::vector<double> ones(nparam, 1.0);
    std::vector<double> hdiag;
    jtj_diagonal(P, L, ones, &hdiag);
    double hmax = 0.0;
    for (double v : hdiag) hmax = std::max(hmax, v);
    double radius = 1.0 / (o.g2o_tau * hmax), nu = 2.0;
    auto grad_max = [&]() { double m = 0.0; for (double g : L.grad) m = std::max(m, std::fabs(g)); return m; };
    double x_cost = L.cost, gmax = grad_max();
    *initial_cost = 2.0 * x_cost;
    IterRecord rec{};
    rec.step_is_successful = 1;
    rec.step_is_valid = 1;
    rec.cost = 2.0 * x_cost;
    rec.gradient_max_norm = gmax;
    rec.trust_region_radius = radius;
    push(rec);
    int term = 1, trial = 0, iter = 0;
    bool stop = false;
    while (!stop && iter < o.max_num_iterations) {
        for (int t = 0;; ++t) {
            rec = IterRecord{};
            rec.iteration = ++trial;
            std::vector<double> y;
            const int lin_it = solve_damped(P, o, po, L, ones, ones, radius, &y);
            rec.linear_solver_iterations = lin_it < 0 ? 0 : lin_it;
            const bool have = lin_it >= 0;
            double mcc = 0.0, cand_cost = std::numeric_limits<double>::max(), rho = -1.0;
            std::vector<double> delta(nparam, 0.0);
            if (have) {
                for (size_t i = 0; i < nparam; ++i) delta[i] = -y[i];
                mcc = model_cost_change(P, L, delta);
                for (size_t i = 0; i < nparam; ++i) cand[i] = x[i] + delta[i];
                for (int c = 0; c < P.nc; ++c) {
                    double* cc = &cand[9 * (size_t)c];
                    const double* xc = &x[9 * (size_t)c];
                    if (P.stereo)
                        se3_left_update(&delta[9 * (size_t)c], xc, xc + 3, cc, cc + 3);
                    else
                        so3_left_update(&delta[9 * (size_t)c], xc, cc);
                }
                if (!evaluate_cost(P, o, cand.data(), cand.data() + ncam, &cand_cost))
                    cand_cost = std::numeric_limits<double>::max();
                rho = 2.0 * (x_cost - cand_cost) / (2.0 * mcc + 1e-3);
                rec.step_norm = norm2(delta.data(), nparam);
                rec.cost_change = 2.0 * (x_cost - cand_cost);
            }
            rec.step_is_valid = have ? 1 : 0;
            rec.model_cost_change = 2.0 * mcc;
            rec.relative_decrease = rho;
            if (have && rho > 0.0 && cand_cost < std::numeric_limits<double>::max()) {
                x = cand;
                linearize(P, o, x.data(), x.data() + ncam, &L);
                x_cost = L.cost;
                gmax = grad_max();
                radius = radius / std::max(1.0 / 3.0, std::min(2.0 / 3.0, 1.0 - std::pow(2.0 * rho - 1.0, 3)));
                nu = 2.0;
                rec.step_is_successful = 1;
                rec.cost = 2.0 * x_cost;
                rec.gradient_max_norm = gmax;
                rec.trust_region_radius = radius;
                push(rec);
                ++iter;
                break;
            }
            radius = radius / nu;
            nu *= 2.0;
            rec.step_is_successful = 0;
            rec.cost = 2.0 * (have ? cand_cost : x_cost);
            rec.gradient_max_norm = gmax;
            rec.trust_region_radius = radius;
            push(rec);
            if (t + 1 >= o.g2o_max_trials || rho == 0.0 || !std::isfinite(radius) || !(radius > 0.0)) {
                stop = true;
                term = 0;
                break;
            }
        }
    }
    std::copy(x.begin(), x.begin() + ncam, cams);
    std::copy(x.begin() + ncam, x.end(), pts);
    *final_cost = 2.0 * x_cost;
    *termination = term;
    if (n_trace) *n_trace = nt;
    return term;
}

// The stereo local BA (g2o mode): cams holds the poses in 9-parameter slots.
int oracle_solve_stereo(int nc, int np, int no, const int* ci, const int* pi, const int* side, const double* obs,
                        double* cams, double* pts, const double* K, const double* ext, const Options* o, IterRecord* trace,
                        int trace_cap, int* n_trace, double* initial_cost, double* final_cost, int* termination) {
    const StereoModel g = make_rig(K, ext, side);
    Problem P{nc, np, no, ci, pi, obs, &g};
    return solve_g2o(P, *o, cams, pts, trace, trace_cap, n_trace, initial_cost, final_cost, termination);
}

int oracle_solve(int nc, int np, int no, const int* ci, const int* pi, const double* obs, double* cams, double* pts,
                 const Options* o, IterRecord* trace, int trace_cap, int* n_trace, double* initial_cost,
                 double* final_cost, int* termination) {
    Problem P{nc, np, no, ci, pi, obs};
    PointObs cam_lists, pt_lists;
    if (g_threads > 1) {
        cam_lists = obs_lists(ci, nc, no);
        pt_lists = obs_lists(pi, np, no);
        P.by_cam = &cam_lists;
        P.by_pt = &pt_lists;
    }
    if (o->mode == 1)
        return solve_g2o(P, *o, cams, pts, trace, trace_cap, n_trace, initial_cost, final_cost, termination);
    const size_t ncam = 9 * (size_t)nc, nparam = ncam + 3 * (size_t)np;
    std::vector<double> x(nparam), cand(nparam);
    std::copy(cams, cams + ncam, x.begin());
    std::copy(pts, pts + 3 * (size_t)np, x.begin() + ncam);
    int nt = 0;
    auto push = [&](const IterRecord& r) { if (trace && nt < trace_cap) trace[nt] = r; ++nt; };
    PointObs po = point_obs(P);

    Linearization L;
    if (!linearize(P, *o, x.data(), x.data() + ncam, &L)) {
        *termination = 2;
        if (n_trace) *n_trace = 0;
        return 2;
    }
    std::vector<double> scale(nparam, 1.0), diag, ones(nparam, 1.0);
    if (o->jacobi_scaling) {
        jtj_diagonal(P, L, ones, &diag);
        for (size_t i = 0; i < nparam; ++i) scale[i] = 1.0 / (1.0 + std::sqrt(diag[i]));
    }
    auto grad_max = [&]() { double m = 0.0; for (double g : L.grad) m = std::max(m, std::fabs(g)); return m; };
    double x_cost = L.cost;
    double radius = o->initial_trust_region_radius, decrease_factor = 2.0;
    double gmax = grad_max();
    *initial_cost = x_cost;
    double min_cost = x_cost;
    std::vector<double> best(x);
    IterRecord rec{};
    rec.iteration = 0;
    rec.step_is_successful = 1;
    rec.step_is_valid = 1;
    rec.cost = x_cost;
    rec.gradient_max_norm = gmax;
    rec.trust_region_radius = radius;
    push(rec);
    int term = -1;
    int iteration = 0, consecutive_invalid = 0;
    bool need_diag = true;
    auto can_continue = [&]() -> bool {
        if (iteration >= o->max_num_iterations) { term = 1; return false; }
        if (gmax <= o->gradient_tolerance) { term = 0; return false; }
        if (radius <= o->min_trust_region_radius) { term = 0; return false; }
        return true;
    };
    while (can_continue()) {
        ++iteration;
        rec = IterRecord{};
        rec.iteration = iteration;
        if (need_diag) {
            jtj_diagonal(P, L, scale, &diag);
            for (size_t i = 0; i < nparam; ++i) diag[i] = std::min(std::max(diag[i], o->min_lm_diagonal), o->max_lm_diagonal);
            need_diag = false;
        }
        std::vector<double> y;
        const int lin_it = solve_damped(P, *o, po, L, scale, diag, radius, &y);
        rec.linear_solver_iterations = lin_it < 0 ? 0 : lin_it;
        std::vector<double> delta(nparam);
        double mcc = 0.0;
        bool valid = false;
        if (lin_it >= 0) {
            for (size_t i = 0; i < nparam; ++i) delta[i] = -y[i] * scale[i];
            mcc = model_cost_change(P, L, delta);
            valid = mcc > 0.0;
        }
        rec.model_cost_change = mcc;
        if (!valid) {
            if (++consecutive_invalid >= o->max_num_consecutive_invalid_steps) { term = 2; break; }
            radius = radius / decrease_factor;
            decrease_factor *= 2.0;
            rec.step_is_valid = 0;
            rec.step_is_successful = 0;
            rec.cost = x_cost;
            rec.gradient_max_norm = gmax;
            rec.trust_region_radius = radius;
            push(rec);
            continue;
        }
        consecutive_invalid = 0;
        rec.step_is_valid = 1;
        for (size_t i = 0; i < nparam; ++i) cand[i] = x[i] + delta[i];
        double cand_cost;
        if (!evaluate_cost(P, *o, cand.data(), cand.data() + ncam, &cand_cost))
            cand_cost = std::numeric_limits<double>::max();
        const double x_norm = norm2(x.data(), nparam);
        const double step_norm = norm2(delta.data(), nparam);
        rec.step_norm = step_norm;
        if (step_norm <= o->parameter_tolerance * (x_norm + o->parameter_tolerance)) { term = 0; break; }
        const double cost_change = x_cost - cand_cost;
        rec.cost_change = cost_change;
        if (std::fabs(cost_change) <= o->function_tolerance * x_cost) { term = 0; break; }
        const double rho = (x_cost - cand_cost) / mcc;
        rec.relative_decrease = rho;
        if (rho > o->min_relative_decrease) {
            x = cand;
            linearize(P, *o, x.data(), x.data() + ncam, &L);
            x_cost = L.cost;
            gmax = grad_max();
            radius = radius / std::max(1.0 / 3.0, 1.0 - std::pow(2.0 * rho - 1.0, 3));
            radius = std::min(o->max_trust_region_radius, radius);
            decrease_factor = 2.0;
            need_diag = true;
            rec.step_is_successful = 1;
            rec.cost = x_cost;
            if (x_cost < min_cost) { min_cost = x_cost; best = x; }
        } else {
            radius = radius / decrease_factor;
            decrease_factor *= 2.0;
            rec.step_is_successful = 0;
            rec.cost = cand_cost;
        }
        rec.gradient_max_norm = gmax;
        rec.trust_region_radius = radius;
        push(rec);
    }
    std::copy(best.begin(), best.begin() + ncam, cams);
    std::copy(best.begin() + ncam, best.end(), pts);
    *final_cost = min_cost;
    *termination = term < 0 ? 1 : term;
    if (n_trace) *n_trace = nt;
    return *termination;
}

// BALProblem.cpp:263-303 on AoS cameras (9 per) and points (3 per).
void oracle_normalize(int nc, int np, double* cams, double* pts) {
    std::vector<double> tmp(np);
    double med[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < np; ++j) tmp[j] = pts[3 * (size_t)j + i];
        med[i] = median(&tmp);
    }
    for (int i = 0; i < np; ++i) {
        const double* p = pts + 3 * (size_t)i;
        tmp[i] = std::fabs(p[0] - med[0]) + std::fabs(p[1] - med[1]) + std::fabs(p[2] - med[2]);
    }
    const double mad = median(&tmp);
    const double scale = 100.0 / mad;
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < 3; ++j) pts[3 * (size_t)i + j] = scale * (pts[3 * (size_t)i + j] - med[j]);
    for (int c = 0; c < nc; ++c) {
        double* cam = cams + 9 * (size_t)c;
        double aa[3], center[3];
        camera_to_aa_center(cam, aa, center);
        for (int j = 0; j < 3; ++j) center[j] = scale * (center[j] - med[j]);
        aa_center_to_camera(aa, center, cam);
    }
}

// BALProblem.cpp:305-336; draws from glibc rand() in the reference's order.
void oracle_perturb(int nc, int np, double* cams, double* pts, double rotation_sigma, double translation_sigma,
                    double point_sigma) {
    if (point_sigma > 0)
        for (int i = 0; i < np; ++i)
            for (int j = 0; j < 3; ++j) pts[3 * (size_t)i + j] += rand_normal() * point_sigma;
    for (int c = 0; c < nc; ++c) {
        double* cam = cams + 9 * (size_t)c;
        double aa[3], center[3];
        camera_to_aa_center(cam, aa, center);
        if (rotation_sigma > 0.0)
            for (int j = 0; j < 3; ++j) aa[j] += rand_normal() * rotation_sigma;
        aa_center_to_camera(aa, center, cam);
        if (translation_sigma > 0.0)
            for (int j = 0; j < 3; ++j) cam[3 + j] += rand_normal() * translation_sigma;
    }
}

void oracle_srand(unsigned seed) { srand(seed); }

}  // extern "C"
