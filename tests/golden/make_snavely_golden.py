"""Generate tests/golden/snavely_golden.json: 50-digit reference values of the Snavely residual and
its 2x9 / 2x3 Jacobians.

The formula is evaluated with mpmath at 50 significant digits, independently of the C++ oracle:
  p = R(w) X + t  (Rodrigues when |w|^2 > DBL_EPSILON, else X + w x X -- the first-order branch of
                   utils/geometry/rotation.h:133-198),
  x' = -p0/p2, y' = -p1/p2, d = 1 + r2 (k1 + k2 r2), pred = f d (x', y')
                  (utils/camera/SnavelyReprojectionError.h:61-87),
  r = pred - obs  (SnavelyReprojectionError.h:49-50).
Derivatives are taken with mpmath.diff at working precision with the branch chosen at the base point
held fixed, which is what forward-mode differentiation of the reference expression computes.

Cases: seeded random cameras/points with theta log-uniform in [1e-3, 3.1], the special angles
{0, 1e-12, 1e-9, just below / at / just above sqrt(DBL_EPSILON), 1e-7, 1e-5, 1e-3, 1, 3.1, 3.14159},
zero distortion, large distortion, and the first observations of data/bal/ladybug49.txt.
Run from the repository root:  python tests/golden/make_snavely_golden.py
"""
import gzip
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
DBL_EPS = mp.mpf(2) ** -52
HERE = os.path.dirname(os.path.abspath(__file__))


def rotate(aa, X, taylor):
    if not taylor:
        th2 = aa[0] ** 2 + aa[1] ** 2 + aa[2] ** 2
        th = mp.sqrt(th2)
        c, s = mp.cos(th), mp.sin(th)
        w = [a / th for a in aa]
        wx = [w[1] * X[2] - w[2] * X[1], w[2] * X[0] - w[0] * X[2], w[0] * X[1] - w[1] * X[0]]
        wd = (w[0] * X[0] + w[1] * X[1] + w[2] * X[2]) * (1 - c)
        return [X[i] * c + wx[i] * s + w[i] * wd for i in range(3)]
    wx = [aa[1] * X[2] - aa[2] * X[1], aa[2] * X[0] - aa[0] * X[2], aa[0] * X[1] - aa[1] * X[0]]
    return [X[i] + wx[i] for i in range(3)]


def residual(params, obs, taylor):
    cam, X = params[:9], params[9:12]
    p = rotate(cam[0:3], X, taylor)
    p = [p[i] + cam[3 + i] for i in range(3)]
    xp, yp = -p[0] / p[2], -p[1] / p[2]
    r2 = xp * xp + yp * yp
    d = 1 + r2 * (cam[7] + cam[8] * r2)
    return [cam[6] * d * xp - obs[0], cam[6] * d * yp - obs[1]]


def case(cam, pt, obs):
    params = [mp.mpf(float(v)) for v in list(cam) + list(pt)]
    ob = [mp.mpf(float(v)) for v in obs]
    th2_double = float(cam[0]) ** 2 + float(cam[1]) ** 2 + float(cam[2]) ** 2  # the reference compares in double
    taylor = not (th2_double > 2.0 ** -52)
    r = residual(params, ob, taylor)
    J = np.zeros((2, 12))
    for j in range(12):
        for row in range(2):
            def f(x, j=j, row=row):
                q = list(params)
                q[j] = x
                return residual(q, ob, taylor)[row]
            J[row, j] = float(mp.diff(f, params[j]))
    return {
        "cam": [float(v) for v in cam],
        "pt": [float(v) for v in pt],
        "obs": [float(v) for v in obs],
        "branch": "taylor" if taylor else "rodrigues",
        "r": [float(v) for v in r],
        "jc": J[:, :9].reshape(-1).tolist(),
        "jp": J[:, 9:].reshape(-1).tolist(),
    }


def random_camera_point(rng, theta, k_scale=1e-3):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    aa = axis * theta
    # choose the point so that it lies in front of the camera (p_z < 0, Snavely convention)
    pc = np.array([rng.normal(0, 2), rng.normal(0, 2), -rng.uniform(4, 30)])
    t = rng.normal(0, 1, size=3)
    # X = R^T (pc - t) with R = exp([aa]x); build R from aa in double for placement only
    th = np.linalg.norm(aa)
    if th > 0:
        k = aa / th
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
    else:
        R = np.eye(3)
    X = R.T @ (pc - t)
    cam = np.concatenate([aa, t, [rng.normal(500, 20), rng.normal(0, k_scale), rng.normal(0, k_scale)]])
    return cam, X


def main():
    rng = np.random.default_rng(20261202)
    cases = []
    for _ in range(120):
        theta = float(np.exp(rng.uniform(np.log(1e-3), np.log(3.1))))
        cam, X = random_camera_point(rng, theta)
        obs = [0.0, 0.0]
        r0 = residual([mp.mpf(float(v)) for v in list(cam) + list(X)], [mp.mpf(0), mp.mpf(0)],
                      not (float(cam[0]) ** 2 + float(cam[1]) ** 2 + float(cam[2]) ** 2 > 2.0 ** -52))
        obs = [float(r0[0]) + rng.normal(0, 0.5), float(r0[1]) + rng.normal(0, 0.5)]
        cases.append(case(cam, X, obs))
    sq = float(mp.sqrt(DBL_EPS))  # 2^-26 exactly
    specials = [0.0, 1e-12, 1e-9, sq * (1 - 2.0 ** -40), sq, sq * (1 + 2.0 ** -40), 1e-7, 1e-5, 1e-3, 1.0, 3.1,
                3.14159]
    for theta in specials:
        for _ in range(3):
            cam, X = random_camera_point(rng, theta)
            if theta == sq:
                cam[0:3] = [sq, 0.0, 0.0]  # theta^2 == DBL_EPSILON exactly: first-order branch
            cases.append(case(cam, X, [rng.normal(0, 100), rng.normal(0, 100)]))
    for _ in range(8):  # zero distortion, and strong distortion so the k1/k2 columns are exercised
        cam, X = random_camera_point(rng, 0.3)
        cam[7:9] = 0.0
        cases.append(case(cam, X, [1.0, -2.0]))
        cam, X = random_camera_point(rng, 0.7, k_scale=0.2)
        cases.append(case(cam, X, [3.0, 4.0]))
    # first observations of the BAL fixture ladybug49 (data/bal/ladybug49.txt)
    with gzip.open(os.path.join(HERE, "ladybug49.txt.gz"), "rt") as fh:
        toks = fh.read().split()
    nc, npt, no = int(toks[0]), int(toks[1]), int(toks[2])
    obs_rows = [toks[3 + 4 * i: 7 + 4 * i] for i in range(no)]
    params = np.array(toks[3 + 4 * no:], dtype=np.float64)
    cams, pts = params[: 9 * nc].reshape(nc, 9), params[9 * nc:].reshape(npt, 3)
    for i in list(range(12)) + [no // 2, no - 1]:
        c, p, x, y = int(obs_rows[i][0]), int(obs_rows[i][1]), float(obs_rows[i][2]), float(obs_rows[i][3])
        cases.append(case(cams[c], pts[p], [x, y]))
    out = {"dps": 50, "generator": "tests/golden/make_snavely_golden.py", "cases": cases}
    with open(os.path.join(HERE, "snavely_golden.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
