"""slamplay_amd -- MI355X-native bundle adjustment for slamplay's BA path.

Python mirror of the host interface (include/mba/bal_problem.h) over the C ABI of
libslamplay_ba.so:

  * BALProblem    -- utils/datasets/BALProblem.{h,cpp}: load a BAL file, Normalize, Perturb,
                     WriteToFile, WriteToPLYFile, and numpy views of the AoS arrays
                     (cameras: 9 doubles each, points: 3, observations: x,y).
  * SolveBA       -- drop-in for SolveBA(BALProblem&) of
                     backend/slam_ba_simple_g2o_ceres/test_bundle_adjustment_ceres.cpp:58-101;
                     solves on the GPU and updates the problem in place.
  * Solver        -- device-resident solver handle (upload once, run many times).
  * eval_residual_jacobian / eval_cost -- per-observation evaluation on the GPU.
  * synthetic     -- seeded synthetic problems of the benchmark configurations.
  * Batch         -- many small BA problems ("windows": the local BA of full_slam/src/backend.cpp:67-199)
                     solved in one launch, one workgroup per window.
  * Server / simulate -- continuous batching of local-BA requests and its Poisson load generator.

Every entry point runs the HIP kernels of libslamplay_ba.so; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _capi as C
from ._capi import BAError

__all__ = ["BALProblem", "SolveBA", "Solver", "Batch", "Server", "simulate", "server_config", "local_ba_windows",
           "default_options", "full_report", "stereo_batch_solve",
           "eval_residual_jacobian", "eval_cost", "synthetic", "synth_params", "BAError", "device_count"]


def default_options(**kw):
    o = C.Options()
    C.lib().ba_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


def device_count():
    n = C.i32(0)
    C.check(C.lib().ba_device_count(ctypes.byref(n)))
    return n.value


class BALProblem:
    """slamplay BALProblem (utils/datasets/BALProblem.h:31-119) backed by a C++ ba_bal handle."""

    def __init__(self, filename=None, use_quaternions=False, _handle=None):
        if _handle is None:
            if filename is None:
                raise ValueError("filename required")
            h = C.vp()
            C.check(C.lib().ba_bal_load(str(filename).encode(), 1 if use_quaternions else 0, ctypes.byref(h)))
            _handle = h
        self._h = _handle
        self._refresh()

    @classmethod
    def from_arrays(cls, camera_index, point_index, observations, cameras, points):
        ci = np.ascontiguousarray(camera_index, dtype=np.int32)
        pi = np.ascontiguousarray(point_index, dtype=np.int32)
        obs = np.ascontiguousarray(observations, dtype=np.float64).reshape(-1)
        cams = np.ascontiguousarray(cameras, dtype=np.float64).reshape(-1)
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
        h = C.vp()
        C.check(C.lib().ba_bal_from_arrays(cams.size // 9, pts.size // 3, ci.size, ci.ctypes.data_as(C.i32p),
                                           pi.ctypes.data_as(C.i32p), obs.ctypes.data_as(C.f64p),
                                           cams.ctypes.data_as(C.f64p), pts.ctypes.data_as(C.f64p), ctypes.byref(h)))
        return cls(_handle=h)

    def _refresh(self):
        v = C.ProblemView()
        C.check(C.lib().ba_bal_view(self._h, ctypes.byref(v)))
        self._view = v

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and C._lib is not None:
            C.lib().ba_bal_destroy(h)
            self._h = None

    # --- accessors (BALProblem.h:55-98) ---
    def camera_block_size(self):
        return self._view.camera_block_size

    def point_block_size(self):
        return 3

    def num_cameras(self):
        return self._view.num_cameras

    def num_points(self):
        return self._view.num_points

    def num_observations(self):
        return self._view.num_observations

    def num_parameters(self):
        return self.camera_block_size() * self.num_cameras() + 3 * self.num_points()

    def _arr(self, ptr, n, dtype):
        if n == 0:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)

    def camera_index(self):
        return self._arr(self._view.camera_index, self.num_observations(), np.int32)

    def point_index(self):
        return self._arr(self._view.point_index, self.num_observations(), np.int32)

    def observations(self):
        return self._arr(self._view.observations, 2 * self.num_observations(), np.float64)

    def mutable_cameras(self):
        return self._arr(self._view.cameras, self.camera_block_size() * self.num_cameras(), np.float64)

    def mutable_points(self):
        return self._arr(self._view.points, 3 * self.num_points(), np.float64)

    cameras = mutable_cameras
    points = mutable_points

    def view(self):
        return self._view

    # --- operations ---
    def Normalize(self):
        C.check(C.lib().ba_bal_normalize(self._h))

    def Perturb(self, rotation_sigma, translation_sigma, point_sigma):
        C.check(C.lib().ba_bal_perturb(self._h, rotation_sigma, translation_sigma, point_sigma))

    def WriteToFile(self, filename, legacy_header=False):
        C.check(C.lib().ba_bal_write(self._h, str(filename).encode(), 1 if legacy_header else 0))

    def WriteToPLYFile(self, filename):
        C.check(C.lib().ba_bal_write_ply(self._h, str(filename).encode()))


def _summary(trace_cap):
    s = C.Summary()
    rows = (C.Iteration * trace_cap)()
    s.trace = ctypes.cast(rows, ctypes.POINTER(C.Iteration))
    s.trace_capacity = trace_cap
    return s, rows


def full_report(summary, num_cameras, num_points, num_observations, options=None):
    """Ceres's Solver::Summary::FullReport() text for a ctypes Summary (what the reference driver
    prints after ceres::Solve, test_bundle_adjustment_ceres.cpp:101)."""
    o = options if options is not None else default_options()
    n = C.i64(0)
    C.check(C.lib().ba_summary_full_report(ctypes.byref(summary), num_cameras, num_points, num_observations,
                                           ctypes.byref(o), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    C.check(C.lib().ba_summary_full_report(ctypes.byref(summary), num_cameras, num_points, num_observations,
                                           ctypes.byref(o), buf, n.value + 1, None))
    return buf.value.decode()


def _summary_dict(s, rows, sizes=None, options=None):
    """The summary as a dict; with sizes = (cameras, points, observations) it carries the
    FullReport text as well."""
    n = min(s.num_iterations, s.trace_capacity)
    trace = [{f: getattr(rows[i], f) for f, _ in C.Iteration._fields_} for i in range(n)]
    out = dict(initial_cost=s.initial_cost, final_cost=s.final_cost, num_iterations=s.num_iterations,
               num_successful_steps=s.num_successful_steps, num_unsuccessful_steps=s.num_unsuccessful_steps,
               termination=s.termination_type, termination_reason=s.termination_reason,
               termination_value=s.termination_value, linear_solver=s.linear_solver_used,
               linear_solver_iterations=s.total_linear_solver_iterations, total_time_ms=s.total_time_ms,
               trace=trace)
    if sizes is not None:
        out["full_report"] = full_report(s, *sizes, options=options)
    return out


def _trace_rows(o):
    """Trace rows a solve can write: one per trial (a g2o iteration may take g2o_max_trials)."""
    trials = o.max_num_iterations * (o.g2o_max_trials if o.mode == 1 else 1)
    return max(8, trials + 2)


def SolveBA(problem, options=None, **kw):
    """Solve `problem` (a BALProblem) on the GPU in place; returns the summary as a dict."""
    o = options if options is not None else default_options(**kw)
    s, rows = _summary(_trace_rows(o))
    C.check(C.lib().ba_solve(ctypes.byref(problem.view()), ctypes.byref(o), ctypes.byref(s)))
    return _summary_dict(s, rows, (problem.num_cameras(), problem.num_points(), problem.num_observations()), o)


class Solver:
    """Device-resident solver: the problem is uploaded once; run() solves from the parameters held
    on the device, reset() restores the initial ones, download() returns (cameras, points)."""

    def __init__(self, problem, options=None, **kw):
        self.options = options if options is not None else default_options(**kw)
        self._nc, self._np = problem.num_cameras(), problem.num_points()
        self._no = problem.num_observations()
        h = C.vp()
        C.check(C.lib().ba_solver_create(ctypes.byref(problem.view()), ctypes.byref(self.options), ctypes.byref(h)))
        self._h = h

    def run(self):
        s, rows = _summary(_trace_rows(self.options))
        C.check(C.lib().ba_solver_run(self._h, ctypes.byref(s)))
        return _summary_dict(s, rows, (self._nc, self._np, self._no), self.options)

    def reset(self):
        C.check(C.lib().ba_solver_reset(self._h))

    def normalize(self):
        """BALProblem::Normalize on the device-held parameters."""
        C.check(C.lib().ba_solver_normalize(self._h))

    def perturb(self, rotation_sigma, translation_sigma, point_sigma, seed=1):
        """BALProblem::Perturb on the device-held parameters, with glibc rand() after srand(seed)."""
        C.check(C.lib().ba_solver_perturb(self._h, seed, rotation_sigma, translation_sigma, point_sigma))

    def download(self):
        cams = np.zeros(9 * self._nc)
        pts = np.zeros(3 * self._np)
        C.check(C.lib().ba_solver_download(self._h, cams.ctypes.data_as(C.f64p), pts.ctypes.data_as(C.f64p)))
        return cams, pts

    def set_profiling(self, on=True):
        C.check(C.lib().ba_solver_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self):
        buf = (C.KernelTime * 128)()
        n = C.i32(0)
        C.check(C.lib().ba_solver_kernel_times(self._h, buf, 128, ctypes.byref(n)))
        return {buf[i].name.decode(): dict(launches=buf[i].launches, total_ms=buf[i].total_ms,
                                           bytes_per_launch=buf[i].bytes_per_launch) for i in range(min(n.value, 128))}

    def info(self):
        b = C.i64(0)
        ls = C.i32(0)
        C.check(C.lib().ba_solver_info(self._h, ctypes.byref(b), ctypes.byref(ls)))
        return dict(device_bytes=b.value, linear_solver=ls.value)

    def close(self):
        if getattr(self, "_h", None) is not None:
            C.lib().ba_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        if C._lib is not None:
            self.close()


def eval_residual_jacobian(problem, precision=C.PRECISION_FP64, jacobians=True):
    """GPU residuals (No,2) and, optionally, Jc (No,2,9) and Jp (No,2,3) in observation order."""
    no = problem.num_observations()
    r = np.zeros((no, 2))
    jc = np.zeros((no, 2, 9)) if jacobians else None
    jp = np.zeros((no, 2, 3)) if jacobians else None
    C.check(C.lib().ba_eval_residual_jacobian(
        ctypes.byref(problem.view()), precision, r.ctypes.data_as(C.f64p),
        jc.ctypes.data_as(C.f64p) if jacobians else None, jp.ctypes.data_as(C.f64p) if jacobians else None))
    return (r, jc, jp) if jacobians else r


def eval_cost(problem, loss=C.LOSS_HUBER, loss_scale=1.0):
    c = C.f64(0)
    C.check(C.lib().ba_eval_cost(ctypes.byref(problem.view()), loss, loss_scale, ctypes.byref(c)))
    return c.value


def synth_params(config, **kw):
    p = C.SynthParams()
    C.check(C.lib().ba_synth_default(config, ctypes.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown synthetic parameter {k}")
        setattr(p, k, v)
    return p


def local_ba_windows(count, seed=1, min_cameras=5, max_cameras=20, min_points=200, max_points=3000):
    """`count` local-BA windows of uneven size (benchmark configuration 4): keyframes uniform on
    [min_cameras, max_cameras], points log-uniform on [min_points, max_points], tracks of 2 frames up to
    the window length, random-walk keyframes, one generator seed per window."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nc = int(rng.integers(min_cameras, max_cameras + 1))
        npts = int(round(np.exp(rng.uniform(np.log(min_points), np.log(max_points)))))
        out.append(synthetic(4, seed=int(seed) * 1_000_003 + i, num_cameras=nc, num_points=npts, max_track=nc))
    return out


def synthetic(config=1, seed=1, ground_truth=False, **kw):
    """Seeded synthetic problem of benchmark configuration `config` (1..5)."""
    p = synth_params(config, seed=seed, **kw)
    h = C.vp()
    if ground_truth:
        # generous capacity: cameras then points
        cap = 9 * p.num_cameras + 3 * max(p.num_points, p.target_observations)
        gt = np.zeros(cap)
        C.check(C.lib().ba_synth_generate(ctypes.byref(p), ctypes.byref(h), gt.ctypes.data_as(C.f64p), cap))
        prob = BALProblem(_handle=h)
        return prob, gt[: prob.num_parameters()]
    C.check(C.lib().ba_synth_generate(ctypes.byref(p), ctypes.byref(h), None, 0))
    return BALProblem(_handle=h)


def _views(problems):
    arr = (C.ProblemView * len(problems))()
    for i, p in enumerate(problems):
        arr[i] = p.view()
    return arr


def _summary_plain(s):
    return dict(initial_cost=s.initial_cost, final_cost=s.final_cost, num_iterations=s.num_iterations,
                num_successful_steps=s.num_successful_steps, num_unsuccessful_steps=s.num_unsuccessful_steps,
                termination=s.termination_type, termination_reason=s.termination_reason,
                termination_value=s.termination_value, linear_solver=s.linear_solver_used,
                finish_ms=s.total_time_ms)


class Batch:
    """Batched local BA: every window (a BALProblem) keeps its own LM state; run() solves all of them
    from the parameters held on the device, download(w) returns window w's (cameras, points)."""

    def __init__(self, problems, options=None, **kw):
        self.options = options if options is not None else default_options(**kw)
        self._shapes = [(p.num_cameras(), p.num_points()) for p in problems]
        views = _views(problems)
        h = C.vp()
        C.check(C.lib().ba_batch_create(views, len(problems), ctypes.byref(self.options), ctypes.byref(h)))
        self._h = h

    def __len__(self):
        return len(self._shapes)

    def run(self):
        sums = (C.Summary * max(1, len(self)))()
        C.check(C.lib().ba_batch_run(self._h, sums))
        return [_summary_plain(sums[i]) for i in range(len(self))]

    def reset(self):
        C.check(C.lib().ba_batch_reset(self._h))

    def download(self, w):
        nc, npt = self._shapes[w]
        cams = np.zeros(9 * nc)
        pts = np.zeros(3 * npt)
        C.check(C.lib().ba_batch_download(self._h, w, cams.ctypes.data_as(C.f64p), pts.ctypes.data_as(C.f64p)))
        return cams, pts

    def set_block(self, threads):
        C.check(C.lib().ba_batch_set_block(self._h, threads))

    def schedule(self):
        """The launch schedule in use (chosen by the autotuner at creation) with its timings."""
        sc = C.BatchSchedule()
        C.check(C.lib().ba_batch_get_schedule(self._h, ctypes.byref(sc)))
        return {f: getattr(sc, f) for f, _ in C.BatchSchedule._fields_}

    def set_schedule(self, schur_threads=None, solve_threads=None, blocks_per_wave=None, lds_cameras=None):
        """Sets the schedule (None everywhere: run the autotuner again)."""
        args = dict(schur_threads=schur_threads, solve_threads=solve_threads, blocks_per_wave=blocks_per_wave,
                    lds_cameras=lds_cameras)
        if all(v is None for v in args.values()):
            C.check(C.lib().ba_batch_set_schedule(self._h, None))
            return
        cur = self.schedule()
        sc = C.BatchSchedule()
        for k, v in args.items():
            setattr(sc, k, v if v is not None else cur[k])
        C.check(C.lib().ba_batch_set_schedule(self._h, ctypes.byref(sc)))

    def info(self):
        t = C.i32(0)
        b = C.i64(0)
        C.check(C.lib().ba_batch_info(self._h, ctypes.byref(t), ctypes.byref(b)))
        return dict(threads=t.value, device_bytes=b.value)

    def close(self):
        if getattr(self, "_h", None) is not None:
            C.lib().ba_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if C._lib is not None:
            self.close()


def server_config(arena_bytes=1 << 30, max_windows=512, iterations_per_step=1, options=None, **kw):
    c = C.ServerConfig()
    c.arena_bytes = arena_bytes
    c.max_windows = max_windows
    c.iterations_per_step = iterations_per_step
    c.options = options if options is not None else default_options(**kw)
    return c


def _result_dict(r, with_params):
    d = dict(tag=r.tag, initial_cost=r.initial_cost, final_cost=r.final_cost, iterations=r.iterations,
             termination=r.termination_type, submit_time_s=r.submit_time_s, admit_time_s=r.admit_time_s,
             finish_time_s=r.finish_time_s, latency_s=r.finish_time_s - r.submit_time_s, evictions=r.evictions)
    if with_params:
        d["cameras"] = np.ctypeslib.as_array(r.cameras, (9 * r.num_cameras,)).copy() if r.num_cameras else np.zeros(0)
        d["points"] = np.ctypeslib.as_array(r.points, (3 * r.num_points,)).copy() if r.num_points else np.zeros(0)
    return d


class Server:
    """Continuous batching of local-BA requests (see ba_server_* in include/mba/ba_c_api.h)."""

    def __init__(self, config=None, **kw):
        self.config = config if config is not None else server_config(**kw)
        h = C.vp()
        C.check(C.lib().ba_server_create(ctypes.byref(self.config), ctypes.byref(h)))
        self._h = h

    def submit(self, problem, tag):
        C.check(C.lib().ba_server_submit(self._h, ctypes.byref(problem.view()), tag))

    def step(self):
        r = C.i32(0)
        q = C.i32(0)
        C.check(C.lib().ba_server_step(self._h, ctypes.byref(r), ctypes.byref(q)))
        return r.value, q.value

    def cancel(self, tag):
        """Request `tag` leaves at the next step with termination BA_USER_FAILURE (4)."""
        C.check(C.lib().ba_server_cancel(self._h, tag))

    def evict(self, tag):
        """Resident window `tag` is moved to host memory at the next step and queued again."""
        C.check(C.lib().ba_server_evict(self._h, tag))

    def poll(self, capacity=256):
        buf = (C.ServerResult * capacity)()
        n = C.i32(0)
        C.check(C.lib().ba_server_poll(self._h, buf, capacity, ctypes.byref(n)))
        return [_result_dict(buf[i], True) for i in range(n.value)]

    def clock(self):
        return C.lib().ba_server_clock(self._h)

    def drain(self, max_steps=1_000_000):
        """Steps until nothing is resident or queued; returns every result."""
        out = []
        for _ in range(max_steps):
            resident, queued = self.step()
            out += self.poll()
            if resident == 0 and queued == 0:
                break
        while True:
            more = self.poll()
            if not more:
                break
            out += more
        return out

    def close(self):
        if getattr(self, "_h", None) is not None:
            C.lib().ba_server_destroy(self._h)
            self._h = None

    def __del__(self):
        if C._lib is not None:
            self.close()


def simulate(pool, rate_hz, count, seed=1, config=None, **kw):
    """Poisson arrivals at rate_hz of `count` requests drawn round-robin from `pool` (BALProblems),
    served by a Server; returns one dict per request (latency_s = finish - scheduled arrival)."""
    cfg = config if config is not None else server_config(**kw)
    views = _views(pool)
    res = (C.ServerResult * max(1, count))()
    C.check(C.lib().ba_server_simulate(ctypes.byref(cfg), views, len(pool), rate_hz, count, seed, res))
    return [_result_dict(res[i], False) for i in range(count)]


# ---------------------------------------------------------------------------------------------
# full_slam's stereo local bundle adjustment (full_slam/src/backend.cpp:67-199)

def stereo_default_options(**kw):
    """ba_stereo_default_options: g2o Levenberg, 10 iterations, Huber 5.991, dense Schur."""
    o = C.Options()
    C.lib().ba_stereo_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


class _StereoArrays:
    """A ba_stereo_view over private copies of one window's arrays (kept alive with the view)."""

    def __init__(self, pose_index, point_index, side, observations, poses, points, K, extrinsics):
        self.pi = np.ascontiguousarray(pose_index, dtype=np.int32)
        self.qi = np.ascontiguousarray(point_index, dtype=np.int32)
        self.sd = np.ascontiguousarray(side, dtype=np.int32)
        self.obs = np.ascontiguousarray(observations, dtype=np.float64).reshape(-1)
        self.P = np.array(poses, dtype=np.float64, copy=True).reshape(-1, 6)
        self.X = np.array(points, dtype=np.float64, copy=True).reshape(-1, 3)
        self.outlier = np.zeros(max(self.pi.size, 1), dtype=np.uint8)
        v = C.StereoView()
        v.num_poses, v.num_points, v.num_observations = self.P.shape[0], self.X.shape[0], self.pi.size
        v.pose_index = self.pi.ctypes.data_as(C.i32p)
        v.point_index = self.qi.ctypes.data_as(C.i32p)
        v.side = self.sd.ctypes.data_as(C.i32p)
        v.observations = self.obs.ctypes.data_as(C.f64p)
        v.poses = self.P.ctypes.data_as(C.f64p)
        v.points = self.X.ctypes.data_as(C.f64p)
        for i in range(4):
            v.K[i] = float(K[i])
        ext = np.asarray(extrinsics, dtype=np.float64).reshape(2, 6)
        for s in range(2):
            for i in range(6):
                v.extrinsics[s][i] = ext[s, i]
        self.view = v


def stereo_batch_solve(windows, options=None, **kw):
    """Solves many full_slam windows together on the batched engine, each as stereo_solve would
    (same model, rules and outlier rule).  windows: dicts with the arguments of stereo_solve
    (pose_index, point_index, side, observations, poses, points, K, extrinsics).  Returns one dict per
    window (poses, points, outlier, chi2_threshold and the summary fields); the inputs are not
    modified."""
    o = options if options is not None else stereo_default_options(**kw)
    keys = ("pose_index", "point_index", "side", "observations", "poses", "points", "K", "extrinsics")
    arrs = [_StereoArrays(*[w[k] for k in keys]) for w in windows]
    n = len(arrs)
    views = (C.StereoView * max(n, 1))()
    for i, a in enumerate(arrs):
        views[i] = a.view
    sums = (C.Summary * max(n, 1))()
    outp = (ctypes.POINTER(ctypes.c_uint8) * max(n, 1))()
    for i, a in enumerate(arrs):
        outp[i] = a.outlier.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    th = np.zeros(max(n, 1))
    C.check(C.lib().ba_stereo_batch_solve(views, n, ctypes.byref(o), sums, outp, th.ctypes.data_as(C.f64p)))
    res = []
    for i, a in enumerate(arrs):
        r = _summary_plain(sums[i])
        r.update(poses=a.P, points=a.X, outlier=a.outlier[: a.pi.size].astype(bool), chi2_threshold=float(th[i]))
        res.append(r)
    return res


def stereo_solve(pose_index, point_index, side, observations, poses, points, K, extrinsics, options=None, **kw):
    """Optimizes the keyframe poses (angle-axis + translation of T, world to body; (n, 6)) and world
    points ((m, 3)) of one local window on the GPU, then applies the reference's outlier rule.
    Returns dict(poses, points, outlier, chi2_threshold, initial_cost, final_cost, trace, ...);
    the inputs are not modified."""
    o = options if options is not None else stereo_default_options(**kw)
    pi_ = np.ascontiguousarray(pose_index, dtype=np.int32)
    qi = np.ascontiguousarray(point_index, dtype=np.int32)
    sd = np.ascontiguousarray(side, dtype=np.int32)
    obs = np.ascontiguousarray(observations, dtype=np.float64).reshape(-1)
    P = np.array(poses, dtype=np.float64, copy=True).reshape(-1, 6)
    X = np.array(points, dtype=np.float64, copy=True).reshape(-1, 3)
    v = C.StereoView()
    v.num_poses, v.num_points, v.num_observations = P.shape[0], X.shape[0], pi_.size
    v.pose_index = pi_.ctypes.data_as(C.i32p)
    v.point_index = qi.ctypes.data_as(C.i32p)
    v.side = sd.ctypes.data_as(C.i32p)
    v.observations = obs.ctypes.data_as(C.f64p)
    v.poses = P.ctypes.data_as(C.f64p)
    v.points = X.ctypes.data_as(C.f64p)
    for i in range(4):
        v.K[i] = float(K[i])
    ext = np.asarray(extrinsics, dtype=np.float64).reshape(2, 6)
    for s in range(2):
        for i in range(6):
            v.extrinsics[s][i] = ext[s, i]
    s, rows = _summary(_trace_rows(o))
    outl = np.zeros(max(pi_.size, 1), dtype=np.uint8)
    th = ctypes.c_double(0.0)
    C.check(C.lib().ba_stereo_solve(ctypes.byref(v), ctypes.byref(o), ctypes.byref(s),
                                    outl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(th)))
    out = _summary_dict(s, rows)
    out.update(poses=P, points=X, outlier=outl[: pi_.size].astype(bool), chi2_threshold=th.value)
    return out


def _aa_to_R(w):
    w = np.asarray(w, dtype=np.float64)
    t = np.linalg.norm(w)
    if t < 1e-12:
        return np.eye(3)
    k = w / t
    Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(t) * Kx + (1 - np.cos(t)) * Kx @ Kx


def _R_to_aa(R):
    c = np.clip((np.trace(R) - 1) / 2, -1.0, 1.0)
    t = np.arccos(c)
    if t < 1e-12:
        return np.zeros(3)
    v = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]]) / (2 * np.sin(t))
    return v * t


def stereo_window(seed=1, num_keyframes=8, num_points=400, baseline=0.5, noise_px=1.0, outlier_fraction=0.02,
                  pose_sigma=(0.01, 0.05), point_sigma=0.1):
    """A synthetic stereo local-BA window of the kind full_slam's backend optimizes: keyframes moving
    forward along z (EuRoC-like K, a horizontal baseline), world points in front of them, left and
    right pixel observations with Gaussian noise and a fraction of gross outliers; poses and points
    perturbed.  Returns dict(pose_index, point_index, side, observations, poses, points, K, extrinsics,
    true_poses, true_points, planted_outlier)."""
    rng = np.random.default_rng(seed)
    K = np.array([458.0, 457.0, 367.0, 248.0])
    ext = np.zeros((2, 6))
    ext[1, 3] = -baseline
    poses, Rs, ts = [], [], []
    for k in range(num_keyframes):
        c = np.array([0.1 * np.sin(0.7 * k), 0.05 * k, 0.3 * k])
        w = rng.normal(0, 0.02, 3)
        R = _aa_to_R(w)
        t = -R @ c
        poses.append(np.concatenate([w, t]))
        Rs.append(R)
        ts.append(t)
    X = np.stack([rng.uniform(-6, 6, num_points), rng.uniform(-4, 4, num_points), rng.uniform(4, 20, num_points)], 1)
    pose_index, point_index, side, obs, bad = [], [], [], [], []
    for p in range(num_points):
        for k in range(num_keyframes):
            q = Rs[k] @ X[p] + ts[k]
            for s in (0, 1):
                pc = q + ext[s, 3:]
                if pc[2] < 0.5:
                    continue
                u, v = K[0] * pc[0] / pc[2] + K[2], K[1] * pc[1] / pc[2] + K[3]
                if not (0 <= u < 752 and 0 <= v < 480):
                    continue
                if rng.random() < outlier_fraction:
                    u, v = u + rng.uniform(-60, 60), v + rng.uniform(-60, 60)
                    bad.append(True)
                else:
                    u, v = u + rng.normal(0, noise_px), v + rng.normal(0, noise_px)
                    bad.append(False)
                pose_index.append(k)
                point_index.append(p)
                side.append(s)
                obs.append((u, v))
    poses = np.array(poses)
    init = poses.copy()
    for k in range(num_keyframes):
        R = _aa_to_R(rng.normal(0, pose_sigma[0], 3)) @ Rs[k]
        init[k, :3] = _R_to_aa(R)
        init[k, 3:] += rng.normal(0, pose_sigma[1], 3)
    return dict(pose_index=np.array(pose_index, np.int32), point_index=np.array(point_index, np.int32),
                side=np.array(side, np.int32), observations=np.array(obs), poses=init,
                points=X + rng.normal(0, point_sigma, X.shape), K=K, extrinsics=ext, true_poses=poses, true_points=X,
                planted_outlier=np.array(bad))
