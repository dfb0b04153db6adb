"""Batched local BA (one workgroup per window) and the serving loop, vs the oracle and vs themselves.

Bars: FP64 windows follow the oracle's dense-Schur LM (final cost within 1e-9 relative, same
termination and step counts on a fixed schedule); mixed-precision windows end within 1e-5 of the
FP64 oracle; a window's result is bitwise the same alone, in any mix and through the server.
"""
import time

import numpy as np
import pytest

import slamplay_amd as S

pytestmark = pytest.mark.gpu

FIXED = dict(function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0)


def window(seed, nc, npts):
    """A local-BA window: random-walk keyframes, tracks of 2..nc frames."""
    return S.synthetic(4, seed=seed, num_cameras=nc, num_points=npts, max_track=nc)


def _same(a, b):
    """Summaries equal in every solver output (finish_ms is a measured time)."""
    strip = lambda d: {k: v for k, v in d.items() if k != "finish_ms"}
    return strip(a) == strip(b)


def _oracle(oracle_mod, p, **kw):
    o = oracle_mod.default_options(linear_solver=1, **kw)
    return oracle_mod.solve(p.num_cameras(), p.num_points(), p.camera_index(), p.point_index(), p.observations(),
                            p.mutable_cameras(), p.mutable_points(), o)


WINDOWS = [(101, 5, 200), (102, 9, 700), (103, 14, 400), (104, 12, 1000), (105, 7, 300)]


def test_fp64_windows_match_oracle(oracle_mod):
    probs = [window(*w) for w in WINDOWS]
    b = S.Batch(probs, precision=0, max_num_iterations=10, **FIXED)
    sums = b.run()
    for w, p in enumerate(probs):
        ref = _oracle(oracle_mod, p, max_num_iterations=10, **FIXED)
        s = sums[w]
        assert s["num_iterations"] == len(ref["trace"])
        assert s["num_successful_steps"] == sum(r["step_is_successful"] for r in ref["trace"][1:])
        assert abs(s["initial_cost"] - ref["trace"][0]["cost"]) <= 1e-12 * ref["trace"][0]["cost"]
        assert abs(s["final_cost"] - ref["final_cost"]) <= 1e-9 * ref["final_cost"], (w, s, ref["final_cost"])
        cams, pts = b.download(w)
        np.testing.assert_allclose(pts, ref["pts"], rtol=0, atol=1e-7 * np.abs(ref["pts"]).max())
        np.testing.assert_allclose(cams, ref["cams"], rtol=0, atol=1e-7 * np.abs(ref["cams"]).max())


def test_fp64_default_tolerances_terminate_like_oracle(oracle_mod):
    probs = [window(*w) for w in WINDOWS[:3]]
    sums = S.Batch(probs, precision=0).run()
    for p, s in zip(probs, sums):
        ref = _oracle(oracle_mod, p)
        assert s["termination"] == ref["termination"]
        assert s["num_iterations"] == len(ref["trace"])
        assert abs(s["final_cost"] - ref["final_cost"]) <= 1e-9 * ref["final_cost"]


def test_mixed_windows_final_cost(oracle_mod):
    probs = [window(*w) for w in WINDOWS] + [window(106, 20, 2500)]
    sums = S.Batch(probs, precision=1, max_num_iterations=10, **FIXED).run()
    for p, s in zip(probs, sums):
        ref = _oracle(oracle_mod, p, max_num_iterations=10, **FIXED)
        assert abs(s["final_cost"] - ref["final_cost"]) <= 1e-5 * ref["final_cost"], (s, ref["final_cost"])


def test_matches_single_problem_solver():
    """The window kernel and the single-problem dense-Schur solver take the same LM steps."""
    probs = [window(*w) for w in WINDOWS[:3]]
    sums = S.Batch(probs, precision=0, max_num_iterations=8, **FIXED).run()
    for p, s in zip(probs, sums):
        q = S.BALProblem.from_arrays(p.camera_index(), p.point_index(), p.observations(), p.mutable_cameras(),
                                     p.mutable_points())
        out = S.SolveBA(q, precision=0, linear_solver=1, max_num_iterations=8, **FIXED)
        assert s["num_iterations"] == out["num_iterations"]
        assert abs(s["final_cost"] - out["final_cost"]) <= 1e-10 * out["final_cost"]


@pytest.mark.parametrize("precision", [0, 1])
def test_alone_and_in_mix_bitwise_identical(precision):
    target = window(201, 10, 600)
    others = [window(202 + i, 5 + (3 * i) % 9, 200 + 150 * i) for i in range(7)]
    kw = dict(precision=precision, max_num_iterations=10)
    alone = S.Batch([target], **kw)
    s_alone = alone.run()[0]
    c_alone, p_alone = alone.download(0)
    mix = S.Batch(others[:3] + [target] + others[3:], **kw)
    s_mix = mix.run()[3]
    c_mix, p_mix = mix.download(3)
    assert _same(s_alone, s_mix)
    np.testing.assert_array_equal(c_alone, c_mix)
    np.testing.assert_array_equal(p_alone, p_mix)
    # every schedule the autotuner may pick gives the same bits too (each block of S is reduced by
    # one wave in pair order, each factor entry by one thread)
    tuned = mix.schedule()
    assert tuned["candidates"] >= 6 and tuned["iteration_ms"] > 0 and tuned["default_iteration_ms"] > 0
    assert tuned["iteration_ms"] <= tuned["default_iteration_ms"]
    sizes = 0
    for ts in (256, 512):
        for tf in (256, 512):
            for bpw, lds in ((1, 21), (2, 0), (4, 10), (8, 21)):
                try:
                    mix.set_schedule(ts, tf, bpw, lds)
                except S.BAError as e:
                    assert e.status == 6  # BA_ERR_CAPACITY
                    continue
                sc = mix.schedule()
                assert sc["blocks_per_wave"] == bpw and sc["lds_cameras"] == lds
                sizes += 1
                mix.reset()
                assert _same(mix.run()[3], s_alone)
                np.testing.assert_array_equal(mix.download(3)[1], p_alone)
    assert sizes >= 8
    # windows compute in FP64 whatever the requested precision
    other = S.Batch([target], **dict(kw, precision=1 - precision))
    assert _same(other.run()[0], s_alone)
    np.testing.assert_array_equal(other.download(0)[1], p_alone)


def test_lanes_bitwise_identical():
    """A batch deals its windows to concurrently running lanes (one per 32 windows, at most four);
    every window's result equals bit for bit that of a one-lane batch holding it."""
    probs = [window(900 + i, 5 + i % 9, 200 + 13 * i) for i in range(128)]
    kw = dict(precision=0, max_num_iterations=6)
    many = S.Batch(probs, **kw)
    assert many.schedule()["lanes"] == 4
    s2 = many.run()
    for off in range(0, 128, 31):
        one = S.Batch(probs[off:off + 31], **kw)
        assert one.schedule()["lanes"] == 1
        s1 = one.run()
        for j in range(len(probs[off:off + 31])):
            w = off + j
            assert _same(s1[j], s2[w])
            c1, p1 = one.download(j)
            c2, p2 = many.download(w)
            np.testing.assert_array_equal(c1, c2)
            np.testing.assert_array_equal(p1, p2)
    assert all(s["finish_ms"] > 0 for s in s2)


def test_reset_rerun_and_empty_window():
    probs = [window(301, 8, 400), S.BALProblem.from_arrays(np.zeros(0, np.int32), np.zeros(0, np.int32),
                                                           np.zeros(0), np.zeros(9 * 2), np.zeros(3 * 4))]
    b = S.Batch(probs, precision=1, max_num_iterations=6)
    a = b.run()
    ca, pa = b.download(0)
    b.reset()
    a2 = b.run()
    assert all(_same(x, y) for x, y in zip(a, a2))
    np.testing.assert_array_equal(ca, b.download(0)[0])
    assert a[1]["termination"] == 0 and a[1]["final_cost"] == 0.0
    info = b.info()
    assert info["threads"] in (256, 512) and info["device_bytes"] > 0


@pytest.mark.parametrize("precision", [0, 1])
def test_large_windows(oracle_mod, precision):
    """30 cameras (and 32 in mixed precision, beyond the 30 an FP32 packed factor fits in LDS): the
    reduced system is factored in the window's global buffer instead of LDS; 36 cameras: beyond the
    batched kernels, solved by the single-problem solver.  All are reported like the others and
    follow the oracle (FP64: 1e-9; mixed: 1e-6 of the FP64 oracle)."""
    probs = [window(401, 30, 800), window(402, 6, 300), window(403, 36, 600)]
    if precision == 1:
        probs.append(window(404, 32, 700))
    sums = S.Batch(probs, precision=precision, max_num_iterations=8, **FIXED).run()
    for p, s in zip(probs, sums):
        ref = _oracle(oracle_mod, p, max_num_iterations=8, **FIXED)
        tol = 1e-9 if precision == 0 else 1e-6
        assert abs(s["final_cost"] - ref["final_cost"]) <= tol * ref["final_cost"], (p.num_cameras(), s)
        if precision == 0:
            assert s["num_iterations"] == len(ref["trace"])


def _with_repeats(p, every=7):
    """p with every `every`-th observation repeated (same camera, point and measurement plus a
    small offset): cameras that observe one point twice."""
    ci, pi, obs = p.camera_index(), p.point_index(), p.observations().reshape(-1, 2)
    k = np.arange(0, len(ci), every)
    obs2 = obs[k] + 0.3
    return S.BALProblem.from_arrays(np.concatenate([ci, ci[k]]).astype(np.int32),
                                    np.concatenate([pi, pi[k]]).astype(np.int32),
                                    np.concatenate([obs, obs2]).reshape(-1), p.mutable_cameras(), p.mutable_points())


@pytest.mark.parametrize("precision", [0, 1])
def test_window_with_repeated_observations(oracle_mod, precision):
    """A camera that observes the same point twice: the cross terms of its diagonal block of S
    (-J~c,x' J~p,x V*^-1 J~p,y' J~c,y, x != y) are formed like the rest; the window follows the
    oracle and the single-problem dense solver (FP64)."""
    p = _with_repeats(window(420, 8, 500))
    s = S.Batch([p], precision=precision, max_num_iterations=8, **FIXED).run()[0]
    ref = _oracle(oracle_mod, p, max_num_iterations=8, **FIXED)
    tol = 1e-9 if precision == 0 else 1e-6
    assert abs(s["final_cost"] - ref["final_cost"]) <= tol * ref["final_cost"], (s, ref["final_cost"])
    # the single-problem dense solver in FP64 (its mixed mode builds S from FP32 Jacobians, a different
    # precision design with its own bar)
    q = S.BALProblem.from_arrays(p.camera_index(), p.point_index(), p.observations(), p.mutable_cameras(),
                                 p.mutable_points())
    out = S.SolveBA(q, precision=0, linear_solver=1, max_num_iterations=8, **FIXED)
    assert abs(s["final_cost"] - out["final_cost"]) <= tol * out["final_cost"]


def test_server_rejects_request_larger_than_arena():
    """A request that needs more than the whole arena ends with a BA_FAILURE result and does not
    block the requests behind it."""
    small = [window(650 + i, 5, 200) for i in range(3)]
    big = window(660, 20, 3000)
    opts = S.default_options(precision=1, max_num_iterations=5)
    srv = S.Server(S.server_config(arena_bytes=8 << 20, max_windows=4, iterations_per_step=1, options=opts))
    srv.submit(big, 0)
    for i, p in enumerate(small):
        srv.submit(p, 1 + i)
    res = {r["tag"]: r for r in srv.drain()}
    assert sorted(res) == [0, 1, 2, 3]
    assert res[0]["termination"] == 2  # BA_FAILURE
    np.testing.assert_array_equal(res[0]["points"], big.mutable_points().reshape(-1))
    for t in (1, 2, 3):
        assert res[t]["termination"] in (0, 1) and res[t]["final_cost"] < res[t]["initial_cost"]


def test_finish_times():
    """Every window reports when it finished within the run; a one-window batch finishes last."""
    probs = [window(450 + i, 5 + i, 200 + 100 * i) for i in range(6)]
    sums = S.Batch(probs, precision=1, max_num_iterations=5).run()
    t = np.array([s["finish_ms"] for s in sums])
    assert (t > 0).all() and np.isfinite(t).all()


def test_invalid_block_size():
    b = S.Batch([window(501, 6, 200)], precision=1)
    with pytest.raises(S.BAError):
        b.set_block(1024)
    with pytest.raises(S.BAError, match="blocks_per_wave"):
        b.set_schedule(256, 256, 3)
    with pytest.raises(S.BAError, match="256 or 512"):
        b.set_schedule(128, 256, 2)
    b.set_schedule()  # autotune again
    assert b.schedule()["candidates"] >= 6


@pytest.mark.parametrize("arena", [1 << 30, 12 << 20])
def test_server_results_equal_alone(arena):
    """Continuous batching: requests admitted while others run (a small arena forces waiting for
    memory) give bitwise the results of solving each window alone."""
    probs = [window(600 + i, 5 + (7 * i) % 16, 200 + 97 * i) for i in range(12)]
    opts = S.default_options(precision=1, max_num_iterations=10)
    srv = S.Server(S.server_config(arena_bytes=arena, max_windows=5, iterations_per_step=1, options=opts))
    for i, p in enumerate(probs[:6]):
        srv.submit(p, i)
    results = []
    for _ in range(3):
        srv.step()
        results += srv.poll()
    for i, p in enumerate(probs[6:]):
        srv.submit(p, 6 + i)
    results += srv.drain()
    assert sorted(r["tag"] for r in results) == list(range(12))
    for r in results:
        alone = S.Batch([probs[r["tag"]]], options=opts)
        s = alone.run()[0]
        cams, pts = alone.download(0)
        assert r["final_cost"] == s["final_cost"] and r["iterations"] == s["num_iterations"]
        np.testing.assert_array_equal(r["cameras"], cams)
        np.testing.assert_array_equal(r["points"], pts)
        assert r["finish_time_s"] >= r["admit_time_s"] >= r["submit_time_s"]


def _alone(p, opts):
    b = S.Batch([p], options=opts)
    s = b.run()[0]
    cams, pts = b.download(0)
    return s, cams, pts


def _step_until_resident(srv, n):
    """Steps until n windows are resident (images are built by the server's threads after submit;
    admission is FIFO and waits for the head's image)."""
    time.sleep(0.5)
    for _ in range(2000):
        resident, _ = srv.step()
        if resident == n:
            return
        time.sleep(0.005)
    raise AssertionError("windows were not admitted")


def test_server_evict_and_readmit_bitwise():
    """A window evicted to host memory mid-solve (twice for one of them) and admitted again into
    another slot and arena range finishes with bitwise the result of a solve never evicted."""
    probs = [window(620 + i, 5 + (5 * i) % 14, 300 + 131 * i) for i in range(6)]
    opts = S.default_options(precision=1, max_num_iterations=10, **FIXED)
    srv = S.Server(S.server_config(arena_bytes=1 << 30, max_windows=3, iterations_per_step=1, options=opts))
    for i, p in enumerate(probs):
        srv.submit(p, i)
    _step_until_resident(srv, 3)  # 0, 1, 2, after at least one iteration each
    srv.step()
    results = srv.poll()
    assert not results
    srv.evict(0)
    srv.evict(2)
    srv.evict(99)  # not resident: ignored
    for _ in range(4):
        srv.step()
        results += srv.poll()
    srv.evict(0)  # a second time if it is resident again by now
    results += srv.drain()
    assert sorted(r["tag"] for r in results) == list(range(6))
    ev = {r["tag"]: r["evictions"] for r in results}
    assert ev[0] >= 1 and ev[2] == 1 and ev[1] == ev[3] == ev[4] == ev[5] == 0
    for r in results:
        s, cams, pts = _alone(probs[r["tag"]], opts)
        assert r["termination"] == s["termination"] and r["termination"] in (0, 1)
        assert r["final_cost"] == s["final_cost"] and r["iterations"] == s["num_iterations"]
        np.testing.assert_array_equal(r["cameras"], cams)
        np.testing.assert_array_equal(r["points"], pts)


def test_server_cancel():
    """Cancelled requests leave with termination BA_USER_FAILURE (4): a resident one with the
    parameters of its last accepted step (their cost is the reported final cost), a queued one with
    the submitted parameters; the other requests are unaffected."""
    probs = [window(640 + i, 6 + 2 * i, 300 + 100 * i) for i in range(5)]
    opts = S.default_options(precision=1, max_num_iterations=10, **FIXED)
    srv = S.Server(S.server_config(arena_bytes=1 << 30, max_windows=2, iterations_per_step=1, options=opts))
    for i, p in enumerate(probs):
        srv.submit(p, i)
    _step_until_resident(srv, 2)  # 0 and 1
    srv.step()
    results = srv.poll()
    assert not results
    srv.cancel(1)    # resident
    srv.cancel(3)    # queued
    srv.cancel(777)  # unknown: ignored
    results += srv.drain()
    res = {r["tag"]: r for r in results}
    assert sorted(res) == [0, 1, 2, 3, 4]
    r1, r3 = res[1], res[3]
    assert r1["termination"] == 4 and r3["termination"] == 4
    assert r1["iterations"] >= 2 and r1["final_cost"] <= r1["initial_cost"]
    q = S.BALProblem.from_arrays(probs[1].camera_index(), probs[1].point_index(), probs[1].observations(),
                                 r1["cameras"].reshape(-1, 9), r1["points"].reshape(-1, 3))
    assert abs(S.eval_cost(q) - r1["final_cost"]) <= 1e-11 * r1["final_cost"]
    assert r3["iterations"] == 0 and np.isnan(r3["final_cost"])
    np.testing.assert_array_equal(r3["cameras"], probs[3].mutable_cameras().reshape(-1))
    np.testing.assert_array_equal(r3["points"], probs[3].mutable_points().reshape(-1))
    for t in (0, 2, 4):
        s, cams, pts = _alone(probs[t], opts)
        assert res[t]["termination"] in (0, 1) and res[t]["final_cost"] == s["final_cost"]
        np.testing.assert_array_equal(res[t]["cameras"], cams)
        np.testing.assert_array_equal(res[t]["points"], pts)


def test_server_large_problems_on_workers():
    """Requests above 32 cameras go to the server's two worker threads, each creating, running and
    destroying a single-problem solver while the resident windows are stepped; their results equal
    bit for bit those of a single-problem solve with the options the workers use (solver choice
    automatic, stream launches), and the windows beside them are unaffected."""
    small = [window(680 + i, 6 + 3 * i, 300 + 150 * i) for i in range(4)]
    large = [window(690, 38, 800), window(691, 40, 900)]
    opts = S.default_options(precision=0, max_num_iterations=6)
    srv = S.Server(S.server_config(arena_bytes=1 << 30, max_windows=4, iterations_per_step=1, options=opts))
    for i, p in enumerate(small + large):
        srv.submit(p, i)
    res = {r["tag"]: r for r in srv.drain()}
    assert sorted(res) == list(range(6))
    for i, p in enumerate(large):
        r = res[4 + i]
        q = S.BALProblem.from_arrays(p.camera_index(), p.point_index(), p.observations(), p.mutable_cameras(),
                                     p.mutable_points())
        out = S.SolveBA(q, precision=0, max_num_iterations=6, linear_solver=0, launch_mode=1)
        assert r["termination"] == out["termination"] and r["iterations"] == out["num_iterations"]
        assert r["final_cost"] == out["final_cost"]
        np.testing.assert_array_equal(r["cameras"], q.mutable_cameras().reshape(-1))
        np.testing.assert_array_equal(r["points"], q.mutable_points().reshape(-1))
    for i, p in enumerate(small):
        s, cams, pts = _alone(p, opts)
        assert res[i]["final_cost"] == s["final_cost"]
        np.testing.assert_array_equal(res[i]["cameras"], cams)
        np.testing.assert_array_equal(res[i]["points"], pts)


def test_simulate_poisson():
    pool = [window(700 + i, 5 + i % 12, 200 + 50 * i) for i in range(16)]
    res = S.simulate(pool, rate_hz=2000.0, count=64, seed=3, max_windows=64, precision=1, max_num_iterations=10)
    assert len(res) == 64
    lat = np.array([r["latency_s"] for r in res])
    assert (lat > 0).all() and np.isfinite(lat).all()
    assert all(r["termination"] in (0, 1) for r in res)
    assert all(r["final_cost"] < r["initial_cost"] for r in res)


def test_mixed_windows_38_final_cost(oracle_mod):
    """38 generated windows (5-20 keyframes) in mixed precision: final costs within 1e-6 of the FP64
    oracle after 10 LM iterations (the FP32 factor preconditions CG against the FP64 system)."""
    wins = S.local_ba_windows(38, seed=11)
    sums = S.Batch(wins, precision=1, max_num_iterations=10, **FIXED).run()
    worst = 0.0
    for p, s in zip(wins, sums):
        ref = _oracle(oracle_mod, p, max_num_iterations=10, **FIXED)
        worst = max(worst, abs(s["final_cost"] - ref["final_cost"]) / ref["final_cost"])
    assert worst < 1e-6, worst
