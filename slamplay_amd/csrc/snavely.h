// snavely.h -- Snavely reprojection residual and its analytic 2x9 / 2x3 Jacobian for the device.
//
// The derivative follows the expression graph of the reference functor evaluated with forward-mode
// dual numbers (utils/camera/SnavelyReprojectionError.h:42-87 on top of
// utils/geometry/rotation.h:133-198), so that inside the DBL_EPSILON first-order branch the result
// is exactly d(X + w x X)/dw = -[X]x and near the branch the rounding behaviour matches autodiff.
//
// Split for MI355X: everything that depends only on the camera (theta, cos, sin, 1/theta, the unit
// axis and their derivatives w.r.t. the angle-axis) is computed once per camera in FP64
// (CameraPrep, ~30 values); the per-observation part carries only the point-dependent terms.
// Kernels walk observations camera by camera, so a camera's prep is uniform across a workgroup.
#pragma once

#include <cfloat>
#include <cmath>

#if defined(__HIPCC__)
#define BA_HD __host__ __device__ __forceinline__
#else
#define BA_HD inline
#endif

namespace mba {

// Per-camera terms of Rodrigues' formula and their derivatives w.r.t. the angle-axis w.
//   taylor == 0: R X = X c + (u x X) s + u (u.X)(1 - c), u = w / theta (rotation.h:147-169)
//   taylor == 1: R X = X + w x X                                       (rotation.h:191-196)
template <typename Real>
struct CameraPrep {
    Real c, s, omc;        // cos(theta), sin(theta), 1 - cos(theta)
    Real u[3];             // unit axis (taylor: the raw angle-axis)
    Real dc[3], ds[3];     // d cos / dw_j, d sin / dw_j  (d(1-c) = -dc)
    Real du[3][3];         // du[j][i] = d u_i / d w_j
    Real t[3];             // translation
    Real f, k1, k2;        // focal length, radial distortion
    double R64[9];         // rotation matrix and translation in FP64: the camera-frame point
    double t64[3];         // p = R X + t of the FP32 path is formed in FP64 (see below)
    int taylor;
};

// Build the prep of one camera from its 9 FP64 parameters (angle-axis, t, f, k1, k2). The branch
// threshold is DBL_EPSILON on theta^2 computed in double, as in the reference for every T.
template <typename Real>
BA_HD void camera_prep(const double cam[9], CameraPrep<Real>& P) {
    const double w0 = cam[0], w1 = cam[1], w2 = cam[2];
    const double th2 = w0 * w0 + w1 * w1 + w2 * w2;
    P.t[0] = (Real)cam[3];
    P.t[1] = (Real)cam[4];
    P.t[2] = (Real)cam[5];
    P.f = (Real)cam[6];
    P.k1 = (Real)cam[7];
    P.k2 = (Real)cam[8];
    for (int i = 0; i < 3; ++i) P.t64[i] = cam[3 + i];
    if (th2 > DBL_EPSILON) {
        const double th = sqrt(th2), c = cos(th), s = sin(th), u[3] = {w0 / th, w1 / th, w2 / th};
        const double K[3][3] = {{0.0, -u[2], u[1]}, {u[2], 0.0, -u[0]}, {-u[1], u[0], 0.0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) P.R64[3 * i + j] = (i == j ? c : 0.0) + s * K[i][j] + (1.0 - c) * u[i] * u[j];
    } else {
        const double K[3][3] = {{0.0, -w2, w1}, {w2, 0.0, -w0}, {-w1, w0, 0.0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) P.R64[3 * i + j] = (i == j ? 1.0 : 0.0) + K[i][j];
    }
    if (th2 > DBL_EPSILON) {
        // Dual-number order: d(theta^2) = 2w, theta = sqrt, d theta = d(theta^2)/(2 theta),
        // d cos = -sin d theta, d sin = cos d theta, d(1/theta) = -d theta / theta^2,
        // u_i = w_i (1/theta) => du_i/dw_j = w_i d(1/theta)_j + delta_ij / theta.
        const double th = sqrt(th2);
        const double two_a_inverse = 1.0 / (2.0 * th);
        const double w[3] = {w0, w1, w2};
        double dth[3];
        for (int j = 0; j < 3; ++j) dth[j] = (w[j] + w[j]) * two_a_inverse;
        const double c = cos(th), s = sin(th);
        const double inv = 1.0 / th;
        const double minus_inv2 = -1.0 / (th * th);
        P.c = (Real)c;
        P.s = (Real)s;
        P.omc = (Real)(1.0 - c);
        for (int j = 0; j < 3; ++j) {
            P.dc[j] = (Real)(-s * dth[j]);
            P.ds[j] = (Real)(c * dth[j]);
        }
        for (int i = 0; i < 3; ++i) P.u[i] = (Real)(w[i] * inv);
        for (int j = 0; j < 3; ++j) {
            const double dinv = dth[j] * minus_inv2;
            for (int i = 0; i < 3; ++i) P.du[j][i] = (Real)(w[i] * dinv + (i == j ? inv : 0.0));
        }
        P.taylor = 0;
    } else {
        P.c = (Real)1;
        P.s = (Real)0;
        P.omc = (Real)0;
        P.u[0] = (Real)w0;
        P.u[1] = (Real)w1;
        P.u[2] = (Real)w2;
        for (int j = 0; j < 3; ++j) {
            P.dc[j] = P.ds[j] = (Real)0;
            for (int i = 0; i < 3; ++i) P.du[j][i] = (Real)0;
        }
        P.taylor = 1;
    }
}

// Camera centre c = -R't of a prep (FP64).
template <typename Real>
BA_HD void camera_centre(const CameraPrep<Real>& P, double c[3]) {
    for (int i = 0; i < 3; ++i) c[i] = -(P.R64[i] * P.t64[0] + P.R64[3 + i] * P.t64[1] + P.R64[6 + i] * P.t64[2]);
}

// Camera-frame point p = R(w) X + t.  FP64: the reference's Rodrigues expression order
// (rotation.h:147-169).  FP32: formed in FP64 from the prep's rotation matrix and rounded once, so
// that points close to the camera center (|p| << |X| + |t|) do not lose their digits to the
// cancellation in R X + t.
template <typename Real>
BA_HD void camera_point(const CameraPrep<Real>& P, const double X[3], Real p[3]) {
    if (sizeof(Real) == sizeof(double)) {
        const Real* u = P.u;
        const Real Xr[3] = {(Real)X[0], (Real)X[1], (Real)X[2]};
        const Real wx[3] = {u[1] * Xr[2] - u[2] * Xr[1], u[2] * Xr[0] - u[0] * Xr[2], u[0] * Xr[1] - u[1] * Xr[0]};
        if (!P.taylor) {
            const Real tmp = (u[0] * Xr[0] + u[1] * Xr[1] + u[2] * Xr[2]) * P.omc;
            for (int i = 0; i < 3; ++i) p[i] = Xr[i] * P.c + wx[i] * P.s + u[i] * tmp;
        } else {
            for (int i = 0; i < 3; ++i) p[i] = Xr[i] + wx[i];
        }
        for (int i = 0; i < 3; ++i) p[i] += P.t[i];
    } else {
        for (int i = 0; i < 3; ++i)
            p[i] = (Real)(P.R64[3 * i] * X[0] + P.R64[3 * i + 1] * X[1] + P.R64[3 * i + 2] * X[2] + P.t64[i]);
    }
}

// Residual only: pred - obs.
template <typename Real>
BA_HD void snavely_residual(const CameraPrep<Real>& P, const double X[3], Real ox, Real oy, Real r[2]) {
    Real p[3];
    camera_point<Real>(P, X, p);
    const Real ginv = (Real)1 / p[2];
    const Real xp = -p[0] * ginv;
    const Real yp = -p[1] * ginv;
    const Real r2 = xp * xp + yp * yp;
    const Real dist = (Real)1 + r2 * (P.k1 + P.k2 * r2);
    const Real fd = P.f * dist;
    r[0] = fd * xp - ox;
    r[1] = fd * yp - oy;
}

// Residual and Jacobians. Jc: 2x9 row-major (w0..2, t0..2, f, k1, k2); Jp: 2x3 row-major (X0..2).
template <typename Real>
BA_HD void snavely_residual_jacobian(const CameraPrep<Real>& P, const double X64[3], Real ox, Real oy, Real r[2],
                                     Real Jc[18], Real Jp[6], bool left = false) {
    const Real* u = P.u;
    const Real X[3] = {(Real)X64[0], (Real)X64[1], (Real)X64[2]};
    Real p[3];
    Real dpw[3][3];  // dpw[i][j] = d p_i / d w_j
    Real dpX[3][3];  // dpX[i][m] = d p_i / d X_m  (= R)
    const Real wx[3] = {u[1] * X[2] - u[2] * X[1], u[2] * X[0] - u[0] * X[2], u[0] * X[1] - u[1] * X[0]};
    if (!P.taylor) {
        const Real dot = u[0] * X[0] + u[1] * X[1] + u[2] * X[2];
        const Real tmp = dot * P.omc;
        for (int j = 0; j < 3; ++j) {
            const Real* du = P.du[j];
            const Real dwx[3] = {du[1] * X[2] - du[2] * X[1], du[2] * X[0] - du[0] * X[2], du[0] * X[1] - du[1] * X[0]};
            const Real ddot = du[0] * X[0] + du[1] * X[1] + du[2] * X[2];
            const Real dtmp = dot * (-P.dc[j]) + ddot * P.omc;
            for (int i = 0; i < 3; ++i)
                dpw[i][j] = X[i] * P.dc[j] + (wx[i] * P.ds[j] + dwx[i] * P.s) + (u[i] * dtmp + du[i] * tmp);
        }
        // d/dX_m: c e_m + (u x e_m) s + u (u_m (1 - c))
        for (int m = 0; m < 3; ++m) {
            const Real dtm = u[m] * P.omc;
            Real uxe[3] = {0, 0, 0};
            if (m == 0) { uxe[1] = u[2]; uxe[2] = -u[1]; }
            if (m == 1) { uxe[0] = -u[2]; uxe[2] = u[0]; }
            if (m == 2) { uxe[0] = u[1]; uxe[1] = -u[0]; }
            for (int i = 0; i < 3; ++i) dpX[i][m] = (i == m ? P.c : (Real)0) + uxe[i] * P.s + u[i] * dtm;
        }
    } else {
        // d(w x X)/dw_j = e_j x X ; d/dX_m = e_m + w x e_m
        dpw[0][0] = 0;     dpw[0][1] = X[2];  dpw[0][2] = -X[1];
        dpw[1][0] = -X[2]; dpw[1][1] = 0;     dpw[1][2] = X[0];
        dpw[2][0] = X[1];  dpw[2][1] = -X[0]; dpw[2][2] = 0;
        dpX[0][0] = 1;     dpX[0][1] = -u[2]; dpX[0][2] = u[1];
        dpX[1][0] = u[2];  dpX[1][1] = 1;     dpX[1][2] = -u[0];
        dpX[2][0] = -u[1]; dpX[2][1] = u[0];  dpX[2][2] = 1;
    }
    camera_point<Real>(P, X64, p);
    if (left) {
        // g2o's camera update R <- exp(d) R (test_bundle_adjustment_g2o.cpp:85): the rotation
        // columns are d p / d d = -[q]x with q = R X, i.e. column j = e_j x q
        double q[3];
        for (int i = 0; i < 3; ++i) q[i] = P.R64[3 * i] * X64[0] + P.R64[3 * i + 1] * X64[1] + P.R64[3 * i + 2] * X64[2];
        dpw[0][0] = 0;           dpw[0][1] = (Real)q[2];  dpw[0][2] = (Real)-q[1];
        dpw[1][0] = (Real)-q[2]; dpw[1][1] = 0;           dpw[1][2] = (Real)q[0];
        dpw[2][0] = (Real)q[1];  dpw[2][1] = (Real)-q[0]; dpw[2][2] = 0;
    }

    // Perspective divide with the dual-number quotient rule: x' = -p0 / p2.
    const Real ginv = (Real)1 / p[2];
    const Real xp = -p[0] * ginv;
    const Real yp = -p[1] * ginv;
    const Real dxp[3] = {-ginv, (Real)0, -xp * ginv};
    const Real dyp[3] = {(Real)0, -ginv, -yp * ginv};
    const Real r2 = xp * xp + yp * yp;
    const Real poly = P.k1 + P.k2 * r2;
    const Real dist = (Real)1 + r2 * poly;
    const Real fd = P.f * dist;
    r[0] = fd * xp - ox;
    r[1] = fd * yp - oy;
    // d pred / d p (2x3)
    Real dpred[2][3];
    for (int m = 0; m < 3; ++m) {
        const Real dr2 = (Real)2 * (xp * dxp[m] + yp * dyp[m]);
        const Real ddist = r2 * (P.k2 * dr2) + dr2 * poly;
        const Real dfd = P.f * ddist;
        dpred[0][m] = fd * dxp[m] + dfd * xp;
        dpred[1][m] = fd * dyp[m] + dfd * yp;
    }
    const Real xy[2] = {xp, yp};
    for (int row = 0; row < 2; ++row) {
        Real* J = Jc + 9 * row;
        for (int j = 0; j < 3; ++j)
            J[j] = dpred[row][0] * dpw[0][j] + dpred[row][1] * dpw[1][j] + dpred[row][2] * dpw[2][j];
        J[3] = dpred[row][0];
        J[4] = dpred[row][1];
        J[5] = dpred[row][2];
        J[6] = dist * xy[row];
        J[7] = P.f * r2 * xy[row];
        J[8] = P.f * (r2 * r2) * xy[row];
        for (int m = 0; m < 3; ++m)
            Jp[3 * row + m] = dpred[row][0] * dpX[0][m] + dpred[row][1] * dpX[1][m] + dpred[row][2] * dpX[2][m];
    }
}

// ---------------------------------------------------------------------------------------------
// Jacobian products without the Jacobian (the PCG matrix-vector products of the Snavely model).
//
// Rodrigues' formula is linear in X, so the rotation columns of d p / d w applied to a direction v
// are dR_v X with the 3x3 matrix dR_v = sum_j v_j dR/dw_j, one per camera and direction.  With
// D = d pred / d p (2x3) and the intrinsic block K (2x3) of snavely_residual_jacobian,
//   Jc v = D (dR_v X + v_t) + K (v_f, v_k1, v_k2),   Jp z = D R z,
//   D a  = -ginv [fd (a0 + x' a2, a1 + y' a2) + fe (x', y') (x' a0 + y' a1 + r2 a2)],
//   D' b = -ginv [fd (b0, b1, s) + fe s (x', y', r2)],  s = x' b0 + y' b1,
//   K (vf, vk1, vk2) = (x', y') (dist vf + f r2 vk1 + f r2^2 vk2),
// with fd = f dist, fe = 2 f (k1 + 2 k2 r2).  The transposed rotation columns sum_obs (d(RX)/dw_j)' g
// become the contractions <dR/dw_j, G> of G = sum_obs g X' (3x3), once per camera.  These are the
// same derivatives as snavely_residual_jacobian's, in about a third of its arithmetic.

// dR_v = sum_j v_j dR/dw_j (row-major).  Differentiating R X = X c + (u x X) s + u (u.X)(1 - c):
//   dR_v = (dc.v) I + (ds.v) [u]x + s [du_v]x + u (omc du_v - (dc.v) u)' + omc du_v u',
// du_v = sum_j v_j du/dw_j; first-order branch R X = X + w x X: dR_v = [v]x.
template <typename Real>
BA_HD void rot_directional(const CameraPrep<Real>& P, const double v[3], double Rd[9]) {
    if (P.taylor) {
        Rd[0] = 0.0;   Rd[1] = -v[2]; Rd[2] = v[1];
        Rd[3] = v[2];  Rd[4] = 0.0;   Rd[5] = -v[0];
        Rd[6] = -v[1]; Rd[7] = v[0];  Rd[8] = 0.0;
        return;
    }
    const double u[3] = {(double)P.u[0], (double)P.u[1], (double)P.u[2]};
    const double s = P.s, omc = P.omc;
    double dcv = 0.0, dsv = 0.0, duv[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 3; ++j) {
        dcv += (double)P.dc[j] * v[j];
        dsv += (double)P.ds[j] * v[j];
        for (int i = 0; i < 3; ++i) duv[i] += v[j] * (double)P.du[j][i];
    }
    const double a[3] = {omc * duv[0] - dcv * u[0], omc * duv[1] - dcv * u[1], omc * duv[2] - dcv * u[2]};
    const double xu[9] = {0.0, -u[2], u[1], u[2], 0.0, -u[0], -u[1], u[0], 0.0};
    const double xd[9] = {0.0, -duv[2], duv[1], duv[2], 0.0, -duv[0], -duv[1], duv[0], 0.0};
    for (int i = 0; i < 3; ++i)
        for (int m = 0; m < 3; ++m)
            Rd[3 * i + m] = (i == m ? dcv : 0.0) + dsv * xu[3 * i + m] + s * xd[3 * i + m] + u[i] * a[m] +
                            omc * duv[i] * u[m];
}

// out_j = <dR/dw_j, G> = sum_im (dR/dw_j)_im G_im for a row-major 3x3 G (the same terms as
// rot_directional, contracted): with gam = (G21 - G12, G02 - G20, G10 - G01),
//   out_j = dc_j (tr G - u'G u) + ds_j u.gam + s du_j.gam + omc (u'G du_j + du_j'G u);
// first-order branch: out = gam.
template <typename Real>
BA_HD void rot_contract(const CameraPrep<Real>& P, const double G[9], double out[3]) {
    const double gam[3] = {G[7] - G[5], G[2] - G[6], G[3] - G[1]};
    if (P.taylor) {
        for (int j = 0; j < 3; ++j) out[j] = gam[j];
        return;
    }
    const double u[3] = {(double)P.u[0], (double)P.u[1], (double)P.u[2]};
    double Gu[3], uG[3];
    for (int i = 0; i < 3; ++i) {
        Gu[i] = G[3 * i] * u[0] + G[3 * i + 1] * u[1] + G[3 * i + 2] * u[2];
        uG[i] = u[0] * G[i] + u[1] * G[3 + i] + u[2] * G[6 + i];
    }
    const double tr_uGu = G[0] + G[4] + G[8] - (u[0] * Gu[0] + u[1] * Gu[1] + u[2] * Gu[2]);
    const double ugam = u[0] * gam[0] + u[1] * gam[1] + u[2] * gam[2];
    const double s = P.s, omc = P.omc;
    for (int j = 0; j < 3; ++j) {
        const double du[3] = {(double)P.du[j][0], (double)P.du[j][1], (double)P.du[j][2]};
        out[j] = (double)P.dc[j] * tr_uGu + (double)P.ds[j] * ugam +
                 s * (du[0] * gam[0] + du[1] * gam[1] + du[2] * gam[2]) +
                 omc * (uG[0] * du[0] + uG[1] * du[1] + uG[2] * du[2] + du[0] * Gu[0] + du[1] * Gu[1] + du[2] * Gu[2]);
    }
}

// The projection's linearization at one observation (camera-frame point formed in FP64 from the
// prep's rotation matrix, then rounded to the working precision).
template <typename Real>
struct ProjLin {
    Real xp, yp, r2, ginv, fd, fe, dist;
};

template <typename Real>
BA_HD void proj_lin(const CameraPrep<Real>& P, const double X[3], ProjLin<Real>& L) {
    Real p[3];
    for (int i = 0; i < 3; ++i)
        p[i] = (Real)(P.R64[3 * i] * X[0] + P.R64[3 * i + 1] * X[1] + P.R64[3 * i + 2] * X[2] + P.t64[i]);
    L.ginv = (Real)1 / p[2];
    L.xp = -p[0] * L.ginv;
    L.yp = -p[1] * L.ginv;
    L.r2 = L.xp * L.xp + L.yp * L.yp;
    const Real poly = P.k1 + P.k2 * L.r2;
    L.dist = (Real)1 + L.r2 * poly;
    L.fd = P.f * L.dist;
    L.fe = (Real)2 * P.f * (poly + P.k2 * L.r2);
}

// a = D q
template <typename Real>
BA_HD void proj_apply(const ProjLin<Real>& L, const Real q[3], Real a[2]) {
    const Real s = L.xp * q[0] + L.yp * q[1] + L.r2 * q[2];
    a[0] = -L.ginv * (L.fd * (q[0] + L.xp * q[2]) + L.fe * L.xp * s);
    a[1] = -L.ginv * (L.fd * (q[1] + L.yp * q[2]) + L.fe * L.yp * s);
}

// g = D' b
template <typename Real>
BA_HD void proj_apply_t(const ProjLin<Real>& L, const Real b[2], Real g[3]) {
    const Real s = L.xp * b[0] + L.yp * b[1];
    const Real fs = L.fe * s;
    g[0] = -L.ginv * (L.fd * b[0] + fs * L.xp);
    g[1] = -L.ginv * (L.fd * b[1] + fs * L.yp);
    g[2] = -L.ginv * (L.fd * s + fs * L.r2);
}

// SO(3) through unit quaternions (w, x, y, z), for the g2o camera update R <- exp(d) R
// (test_bundle_adjustment_g2o.cpp:85) and its angle-axis read-back rotation.log() (:57-63).
// exp: q = (cos(t/2), sin(t/2) w / t), t = |w|, with the series of cos(t/2) and sin(t/2)/t below
// t = 1e-10; log: w = 2 atan(n / q_w) v / n with n = |v|, the series 2 / q_w - 2 n^2 / (3 q_w^3)
// below n = 1e-10 and +-pi / n when |q_w| < 1e-10 (the conventions of Sophus's SO3 exp/log).
BA_HD void so3_exp_quat(const double w[3], double q[4]) {
    const double t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double t = sqrt(t2);
    double re, im;
    if (t < 1e-10) {
        const double t4 = t2 * t2;
        re = 1.0 - t2 / 8.0 + t4 / 384.0;
        im = 0.5 - t2 / 48.0 + t4 / 3840.0;
    } else {
        re = cos(0.5 * t);
        im = sin(0.5 * t) / t;
    }
    q[0] = re;
    q[1] = im * w[0];
    q[2] = im * w[1];
    q[3] = im * w[2];
}

BA_HD void quat_mul(const double a[4], const double b[4], double c[4]) {
    c[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    c[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    c[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    c[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
}

BA_HD void so3_log_quat(const double q[4], double w[3]) {
    const double n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const double n = sqrt(n2), qw = q[0];
    double k;
    if (n < 1e-10) {
        k = 2.0 / qw - 2.0 * n2 / (3.0 * qw * qw * qw);
    } else if (fabs(qw) < 1e-10) {
        k = (qw >= 0.0 ? 3.14159265358979323846 : -3.14159265358979323846) / n;
    } else {
        k = 2.0 * atan(n / qw) / n;
    }
    w[0] = k * q[1];
    w[1] = k * q[2];
    w[2] = k * q[3];
}

// angle-axis of exp(d) exp(w)
BA_HD void so3_left_update(const double d[3], const double w[3], double out[3]) {
    double qd[4], qw[4], qc[4];
    so3_exp_quat(d, qd);
    so3_exp_quat(w, qw);
    quat_mul(qd, qw, qc);
    so3_log_quat(qc, out);
}

// Huber corrector weight sqrt(rho'(s)) and the cost term rho(s)/2 (ceres::HuberLoss(a) with the
// residual corrector for rho'' <= 0).  loss == 0: no loss.
template <typename Real>
BA_HD Real huber_sqrt_rho1(Real s, int loss, Real a) {
    if (loss == 1 && s > a * a) {
        const Real rho1 = a / sqrt(s);
        const Real tiny = sizeof(Real) == 8 ? (Real)DBL_MIN : (Real)FLT_MIN;
        return sqrt(rho1 > tiny ? rho1 : tiny);
    }
    return (Real)1;
}

BA_HD double half_rho(double s, int loss, double a) {
    if (loss == 1 && s > a * a) return 0.5 * (2.0 * a * sqrt(s) - a * a);
    return 0.5 * s;
}

}  // namespace mba
