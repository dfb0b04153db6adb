// stereo.h -- the residual of the full_slam local bundle adjustment (a second camera model).
//
// full_slam/include/myslam/g2o_types.h:122-175 (EdgeProjection) and full_slam/src/backend.cpp:67-199:
// a keyframe pose T (SE3, world to body) observes world point X through the left or the right
// camera of a calibrated stereo rig (extrinsic E_s, intrinsics K):
//   p = E_s (T X),   pred = (fx p_x / p_z + cx, fy p_y / p_z + cy),   e = measurement - pred.
// The pose is updated by left multiplication T <- exp(xi) T (g2o_types.h:56-62), xi = (rho, phi)
// with the translation part first (Sophus's SE3 tangent order).  The edge's own Jacobian
// (g2o_types.h:147-169) is used as written, with X, Y, Z = p: it is the derivative with respect to
// a left perturbation of the composite pose E_s T, which equals that of T for the left camera
// (E = identity) and is the reference's choice for the right one.
//
// Poses are stored in the 9-parameter camera slots of the solver: angle-axis of R, t, then three
// unused entries whose Jacobian columns are zero (their step is zero under any positive damping).
#pragma once

#include "snavely.h"

namespace mba {

struct StereoRig {
    double K[4];      // fx, fy, cx, cy
    double R[2][9];   // extrinsic rotations (row-major) of the left and right cameras
    double t[2][3];   // extrinsic translations
};

// camera-frame point of side s: p = E_s (R X + t), in FP64
template <typename Real>
BA_HD void stereo_point(const CameraPrep<Real>& P, const double X[3], const StereoRig& g, int s, double p[3]) {
    double q[3];
    for (int i = 0; i < 3; ++i) q[i] = P.R64[3 * i] * X[0] + P.R64[3 * i + 1] * X[1] + P.R64[3 * i + 2] * X[2] + P.t64[i];
    for (int i = 0; i < 3; ++i) p[i] = g.R[s][3 * i] * q[0] + g.R[s][3 * i + 1] * q[1] + g.R[s][3 * i + 2] * q[2] + g.t[s][i];
}

template <typename Real>
BA_HD void stereo_residual(const CameraPrep<Real>& P, const double X[3], const StereoRig& g, int s, double ou, double ov,
                           double r[2]) {
    double p[3];
    stereo_point<Real>(P, X, g, s, p);
    r[0] = ou - (g.K[0] * p[0] / p[2] + g.K[2]);
    r[1] = ov - (g.K[1] * p[1] / p[2] + g.K[3]);
}

// Residual, pose Jacobian (columns 0-5 of Jc: rho then phi; 6-8 zero) and point Jacobian
// Jp = J_rho E_s.R R (g2o_types.h:171-172).
template <typename Real>
BA_HD void stereo_residual_jacobian(const CameraPrep<Real>& P, const double X64[3], const StereoRig& g, int s, Real ou,
                                    Real ov, Real r[2], Real Jc[18], Real Jp[6]) {
    double p[3];
    stereo_point<Real>(P, X64, g, s, p);
    const Real fx = (Real)g.K[0], fy = (Real)g.K[1];
    const Real x = (Real)p[0], y = (Real)p[1], z = (Real)p[2];
    r[0] = ou - (fx * x / z + (Real)g.K[2]);
    r[1] = ov - (fy * y / z + (Real)g.K[3]);
    const Real zi = (Real)1 / (z + (Real)1e-18), zi2 = zi * zi;
    Real* J0 = Jc;
    Real* J1 = Jc + 9;
    J0[0] = -fx * zi;
    J0[1] = 0;
    J0[2] = fx * x * zi2;
    J0[3] = fx * x * y * zi2;
    J0[4] = -fx - fx * x * x * zi2;
    J0[5] = fx * y * zi;
    J1[0] = 0;
    J1[1] = -fy * zi;
    J1[2] = fy * y * zi2;
    J1[3] = fy + fy * y * y * zi2;
    J1[4] = -fy * x * y * zi2;
    J1[5] = -fy * x * zi;
    J0[6] = J0[7] = J0[8] = J1[6] = J1[7] = J1[8] = 0;
    // E_s.R * R, in FP64
    double ER[9];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k)
            ER[3 * i + k] = g.R[s][3 * i] * P.R64[k] + g.R[s][3 * i + 1] * P.R64[3 + k] + g.R[s][3 * i + 2] * P.R64[6 + k];
    for (int row = 0; row < 2; ++row)
        for (int k = 0; k < 3; ++k)
            Jp[3 * row + k] = Jc[9 * row] * (Real)ER[k] + Jc[9 * row + 1] * (Real)ER[3 + k] + Jc[9 * row + 2] * (Real)ER[6 + k];
}

// T <- exp(xi) T for the pose stored as angle-axis w and translation t: R' = exp(phi) R and
// t' = exp(phi) t + V(phi) rho, V the SO(3) left Jacobian I + (1 - cos a)/a^2 [phi]x +
// (a - sin a)/a^3 [phi]x^2 (I + [phi]x / 2 below a = 1e-10), as Sophus's SE3::exp defines it.
BA_HD void se3_left_update(const double xi[6], const double w[3], const double t[3], double w_out[3], double t_out[3]) {
    const double* rho = xi;
    const double* phi = xi + 3;
    so3_left_update(phi, w, w_out);
    double q[4];
    so3_exp_quat(phi, q);
    // rotate t by the unit quaternion q: t + 2 q_w (v x t) + 2 v x (v x t)
    const double v[3] = {q[1], q[2], q[3]};
    const double c1[3] = {v[1] * t[2] - v[2] * t[1], v[2] * t[0] - v[0] * t[2], v[0] * t[1] - v[1] * t[0]};
    const double c2[3] = {v[1] * c1[2] - v[2] * c1[1], v[2] * c1[0] - v[0] * c1[2], v[0] * c1[1] - v[1] * c1[0]};
    const double a2 = phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2];
    const double a = sqrt(a2);
    // V = I + k1 [phi]x + k2 [phi]x^2 with k1 = (1 - cos a)/a^2 written as 2 sin^2(a/2)/a^2 and
    // k2 = (a - sin a)/a^3 by its Taylor series below 0.01 (both free of cancellation)
    const double sh = sin(0.5 * a);
    const double k1 = a < 1e-150 ? 0.5 : 2.0 * sh * sh / a2;
    const double k2 = a < 1e-2 ? 1.0 / 6.0 - a2 / 120.0 + a2 * a2 / 5040.0 : (a - sin(a)) / (a2 * a);
    const double pr[3] = {phi[1] * rho[2] - phi[2] * rho[1], phi[2] * rho[0] - phi[0] * rho[2], phi[0] * rho[1] - phi[1] * rho[0]};
    const double ppr[3] = {phi[1] * pr[2] - phi[2] * pr[1], phi[2] * pr[0] - phi[0] * pr[2], phi[0] * pr[1] - phi[1] * pr[0]};
    for (int i = 0; i < 3; ++i)
        t_out[i] = (t[i] + 2.0 * q[0] * c1[i] + 2.0 * c2[i]) + (rho[i] + k1 * pr[i] + k2 * ppr[i]);
}

}  // namespace mba
