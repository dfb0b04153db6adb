"""ctypes binding of the CPU FP64 oracle (oracle/ba_oracle.cpp).

Test infrastructure only: imported by tests/, bench.py's cpu_baseline leg and
__graft_entry__.smoke() as the checker.  The product package never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SLAMPLAY_ORACLE_LIB selects another build of the same source (e.g. `make -C oracle asan`)
LIB_PATH = os.environ.get("SLAMPLAY_ORACLE_LIB") or os.path.join(_HERE, "liboracle_ba.so")

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class Options(ctypes.Structure):
    """Mirror of oracle::Options (Ceres Solver::Options defaults unless stated)."""

    _fields_ = [
        ("max_num_iterations", ctypes.c_int32),
        ("max_num_consecutive_invalid_steps", ctypes.c_int32),
        ("jacobi_scaling", ctypes.c_int32),
        ("loss", ctypes.c_int32),
        ("linear_solver", ctypes.c_int32),
        ("pcg_max_iterations", ctypes.c_int32),
        ("loss_scale", ctypes.c_double),
        ("function_tolerance", ctypes.c_double),
        ("gradient_tolerance", ctypes.c_double),
        ("parameter_tolerance", ctypes.c_double),
        ("initial_trust_region_radius", ctypes.c_double),
        ("max_trust_region_radius", ctypes.c_double),
        ("min_trust_region_radius", ctypes.c_double),
        ("min_relative_decrease", ctypes.c_double),
        ("min_lm_diagonal", ctypes.c_double),
        ("max_lm_diagonal", ctypes.c_double),
        ("pcg_tolerance", ctypes.c_double),
        ("mode", ctypes.c_int32),
        ("g2o_max_trials", ctypes.c_int32),
        ("g2o_tau", ctypes.c_double),
    ]


class IterRecord(ctypes.Structure):
    _fields_ = [
        ("iteration", ctypes.c_int32),
        ("step_is_successful", ctypes.c_int32),
        ("step_is_valid", ctypes.c_int32),
        ("linear_solver_iterations", ctypes.c_int32),
        ("cost", ctypes.c_double),
        ("cost_change", ctypes.c_double),
        ("gradient_max_norm", ctypes.c_double),
        ("step_norm", ctypes.c_double),
        ("relative_decrease", ctypes.c_double),
        ("trust_region_radius", ctypes.c_double),
        ("model_cost_change", ctypes.c_double),
    ]


def default_options(**kw):
    o = Options(
        max_num_iterations=50,
        max_num_consecutive_invalid_steps=5,
        jacobi_scaling=1,
        loss=1,
        linear_solver=1,
        pcg_max_iterations=50,
        loss_scale=1.0,
        function_tolerance=1e-6,
        gradient_tolerance=1e-10,
        parameter_tolerance=1e-8,
        initial_trust_region_radius=1e4,
        max_trust_region_radius=1e16,
        min_trust_region_radius=1e-32,
        min_relative_decrease=1e-3,
        min_lm_diagonal=1e-6,
        max_lm_diagonal=1e32,
        pcg_tolerance=1e-6,
        mode=0,
        g2o_max_trials=10,
        g2o_tau=1e-5,
    )
    for k, v in kw.items():
        setattr(o, k, v)
    return o


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_eval.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.oracle_residual.argtypes = [_f64p, _f64p, ctypes.c_double, ctypes.c_double, _f64p]
        L.oracle_rotate_point.argtypes = [_f64p, _f64p, _f64p]
        L.oracle_jacobian_left.argtypes = [_f64p, _f64p, ctypes.c_double, ctypes.c_double, _f64p, _f64p, _f64p]
        L.oracle_so3_left_update.argtypes = [_f64p, _f64p, _f64p]
        L.oracle_se3_left_update.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p]
        if hasattr(L, "oracle_set_threads"):
            L.oracle_set_threads.argtypes = [ctypes.c_int]
            L.oracle_set_threads.restype = ctypes.c_int
        L.oracle_stereo_residual_jacobian.argtypes = [_f64p, _f64p, _f64p, _f64p, ctypes.c_int, ctypes.c_double,
                                                      ctypes.c_double, _f64p, _f64p, _f64p]
        L.oracle_stereo_residual.argtypes = [_f64p, _f64p, _f64p, _f64p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             _f64p]
        L.oracle_solve_stereo.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, _f64p, _f64p,
                                          _f64p, _f64p, _f64p, ctypes.POINTER(Options), ctypes.POINTER(IterRecord),
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_int), _f64p, _f64p,
                                          ctypes.POINTER(ctypes.c_int)]
        L.oracle_huber.argtypes = [ctypes.c_double, ctypes.c_double, _f64p]
        L.oracle_cost.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, ctypes.c_int, ctypes.c_double]
        L.oracle_cost.restype = ctypes.c_double
        L.oracle_lm_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                     ctypes.POINTER(Options), ctypes.c_double, _f64p, _f64p, _f64p]
        L.oracle_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                   ctypes.POINTER(Options), ctypes.POINTER(IterRecord), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), _f64p, _f64p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_normalize.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, _f64p]
        L.oracle_perturb.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, _f64p, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double]
        L.oracle_srand.argtypes = [ctypes.c_uint]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def eval_residual_jacobian(cam_idx, pt_idx, obs, cams, pts):
    """Raw residuals (No,2), Jc (No,2,9), Jp (No,2,3) in observation order."""
    ci, pi = _c(cam_idx, np.int32), _c(pt_idx, np.int32)
    obs, cams, pts = _c(obs, np.float64), _c(cams, np.float64), _c(pts, np.float64)
    no = ci.size
    r = np.zeros((no, 2))
    jc = np.zeros((no, 2, 9))
    jp = np.zeros((no, 2, 3))
    lib().oracle_eval(no, _p(ci, _i32p), _p(pi, _i32p), _p(obs, _f64p), _p(cams, _f64p), _p(pts, _f64p),
                      _p(r, _f64p), _p(jc, _f64p), _p(jp, _f64p))
    return r, jc, jp


def residual(cam, pt, ox, oy):
    cam, pt = _c(cam, np.float64), _c(pt, np.float64)
    r = np.zeros(2)
    lib().oracle_residual(_p(cam, _f64p), _p(pt, _f64p), ox, oy, _p(r, _f64p))
    return r


def rotate_point(aa, pt):
    aa, pt = _c(aa, np.float64), _c(pt, np.float64)
    out = np.zeros(3)
    lib().oracle_rotate_point(_p(aa, _f64p), _p(pt, _f64p), _p(out, _f64p))
    return out


def jacobian_left(cam, pt, ox, oy):
    """g2o-mode residual (2,), Jc (2,9) with rotation columns for exp(d) R, Jp (2,3)."""
    cam, pt = _c(cam, np.float64), _c(pt, np.float64)
    r, jc, jp = np.zeros(2), np.zeros((2, 9)), np.zeros((2, 3))
    lib().oracle_jacobian_left(_p(cam, _f64p), _p(pt, _f64p), ox, oy, _p(r, _f64p), _p(jc, _f64p), _p(jp, _f64p))
    return r, jc, jp


def so3_left_update(d, w):
    """angle-axis of exp(d) exp(w)."""
    d, w = _c(d, np.float64), _c(w, np.float64)
    out = np.zeros(3)
    lib().oracle_so3_left_update(_p(d, _f64p), _p(w, _f64p), _p(out, _f64p))
    return out


def set_threads(n):
    """Threads for solve() (a timing aid: results are bitwise the same for any count); returns the
    count in use (1 when the library was built without OpenMP)."""
    return lib().oracle_set_threads(int(n))


def se3_left_update(xi, w, t):
    """(angle-axis, translation) of exp(xi) T for T = (w, t), xi = (rho, phi)."""
    xi, w, t = _c(xi, np.float64), _c(w, np.float64), _c(t, np.float64)
    wo, to = np.zeros(3), np.zeros(3)
    lib().oracle_se3_left_update(_p(xi, _f64p), _p(w, _f64p), _p(t, _f64p), _p(wo, _f64p), _p(to, _f64p))
    return wo, to


def stereo_residual_jacobian(pose6, pt, K, ext, side, ou, ov):
    """full_slam edge: residual (2,), pose Jacobian (2,6), point Jacobian (2,3)."""
    cam = np.zeros(9)
    cam[:6] = pose6
    pt, K, ext = _c(pt, np.float64), _c(K, np.float64), _c(ext, np.float64)
    r, jc, jp = np.zeros(2), np.zeros((2, 9)), np.zeros((2, 3))
    lib().oracle_stereo_residual_jacobian(_p(cam, _f64p), _p(pt, _f64p), _p(K, _f64p), _p(ext, _f64p), int(side),
                                          float(ou), float(ov), _p(r, _f64p), _p(jc, _f64p), _p(jp, _f64p))
    return r, jc[:, :6], jp


def stereo_residual(pose6, pt, K, ext, side, ou, ov):
    cam = np.zeros(9)
    cam[:6] = pose6
    pt, K, ext = _c(pt, np.float64), _c(K, np.float64), _c(ext, np.float64)
    r = np.zeros(2)
    lib().oracle_stereo_residual(_p(cam, _f64p), _p(pt, _f64p), _p(K, _f64p), _p(ext, _f64p), int(side), float(ou),
                                 float(ov), _p(r, _f64p))
    return r


def solve_stereo(pose_index, point_index, side, obs, poses, pts, K, ext, options, trace_cap=512):
    """g2o-mode LM on the full_slam stereo model; poses (n, 6).  Returns dict like solve()."""
    ci, pi, sd = _c(pose_index, np.int32), _c(point_index, np.int32), _c(side, np.int32)
    obs = _c(obs, np.float64).reshape(-1)
    P = np.asarray(poses, dtype=np.float64).reshape(-1, 6)
    cams = np.zeros((P.shape[0], 9))
    cams[:, :6] = P
    cams = cams.reshape(-1)
    pts = np.array(pts, dtype=np.float64, copy=True).reshape(-1)
    K, ext = _c(K, np.float64), _c(ext, np.float64).reshape(-1)
    trace = (IterRecord * trace_cap)()
    nt, term = ctypes.c_int(0), ctypes.c_int(0)
    ic, fc = np.zeros(1), np.zeros(1)
    lib().oracle_solve_stereo(P.shape[0], pts.size // 3, ci.size, _p(ci, _i32p), _p(pi, _i32p), _p(sd, _i32p),
                              _p(obs, _f64p), _p(cams, _f64p), _p(pts, _f64p), _p(K, _f64p), _p(ext, _f64p),
                              ctypes.byref(options), trace, trace_cap, ctypes.byref(nt), _p(ic, _f64p), _p(fc, _f64p),
                              ctypes.byref(term))
    recs = trace[: min(nt.value, trace_cap)]
    tr = [{f: getattr(r, f) for f, _ in IterRecord._fields_} for r in recs]
    return dict(poses=cams.reshape(-1, 9)[:, :6].copy(), pts=pts.reshape(-1, 3), initial_cost=ic[0], final_cost=fc[0],
                termination=term.value, trace=tr)


def huber(s, a=1.0):
    rho = np.zeros(3)
    lib().oracle_huber(float(s), float(a), _p(rho, _f64p))
    return rho


def cost(cam_idx, pt_idx, obs, cams, pts, loss=1, loss_scale=1.0):
    ci, pi = _c(cam_idx, np.int32), _c(pt_idx, np.int32)
    obs, cams, pts = _c(obs, np.float64), _c(cams, np.float64), _c(pts, np.float64)
    return lib().oracle_cost(ci.size, _p(ci, _i32p), _p(pi, _i32p), _p(obs, _f64p), _p(cams, _f64p),
                             _p(pts, _f64p), loss, loss_scale)


def lm_step(nc, np_, cam_idx, pt_idx, obs, cams, pts, radius, options=None):
    """One damped LM step at x (scaling from x). Returns (delta, model_cost_change, scale, iters)."""
    o = options or default_options()
    ci, pi = _c(cam_idx, np.int32), _c(pt_idx, np.int32)
    obs, cams, pts = _c(obs, np.float64), _c(cams, np.float64), _c(pts, np.float64)
    n = 9 * nc + 3 * np_
    delta = np.zeros(n)
    scale = np.zeros(n)
    mcc = ctypes.c_double(0.0)
    it = lib().oracle_lm_step(nc, np_, ci.size, _p(ci, _i32p), _p(pi, _i32p), _p(obs, _f64p), _p(cams, _f64p),
                              _p(pts, _f64p), ctypes.byref(o), radius, _p(delta, _f64p), ctypes.byref(mcc),
                              _p(scale, _f64p))
    return delta, mcc.value, scale, it


def solve(nc, np_, cam_idx, pt_idx, obs, cams, pts, options=None, trace_cap=512):
    """Ceres-style LM. Returns dict(cams, pts, initial_cost, final_cost, termination, trace)."""
    o = options or default_options()
    ci, pi = _c(cam_idx, np.int32), _c(pt_idx, np.int32)
    obs = _c(obs, np.float64)
    cams = np.array(cams, dtype=np.float64, copy=True).reshape(-1)
    pts = np.array(pts, dtype=np.float64, copy=True).reshape(-1)
    trace = (IterRecord * trace_cap)()
    nt = ctypes.c_int(0)
    ic, fc = ctypes.c_double(0), ctypes.c_double(0)
    term = ctypes.c_int(0)
    lib().oracle_solve(nc, np_, ci.size, _p(ci, _i32p), _p(pi, _i32p), _p(obs, _f64p), _p(cams, _f64p),
                       _p(pts, _f64p), ctypes.byref(o), trace, trace_cap, ctypes.byref(nt), ctypes.byref(ic),
                       ctypes.byref(fc), ctypes.byref(term))
    recs = [trace[i] for i in range(min(nt.value, trace_cap))]
    tr = [{f: getattr(r, f) for f, _ in IterRecord._fields_} for r in recs]
    return dict(cams=cams, pts=pts, initial_cost=ic.value, final_cost=fc.value, termination=term.value, trace=tr)


def normalize(nc, np_, cams, pts):
    cams = np.array(cams, dtype=np.float64, copy=True).reshape(-1)
    pts = np.array(pts, dtype=np.float64, copy=True).reshape(-1)
    lib().oracle_normalize(nc, np_, _p(cams, _f64p), _p(pts, _f64p))
    return cams, pts


def perturb(nc, np_, cams, pts, rotation_sigma, translation_sigma, point_sigma, seed=1):
    """Reference Perturb on glibc rand(), seeded with srand(seed) first (glibc's default is 1)."""
    cams = np.array(cams, dtype=np.float64, copy=True).reshape(-1)
    pts = np.array(pts, dtype=np.float64, copy=True).reshape(-1)
    lib().oracle_srand(seed)
    lib().oracle_perturb(nc, np_, _p(cams, _f64p), _p(pts, _f64p), rotation_sigma, translation_sigma, point_sigma)
    return cams, pts
