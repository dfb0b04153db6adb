"""Host layer of the engine (no GPU needed): BAL I/O, Normalize/Perturb, synthetic problems, the
C ABI surface and its argument validation."""
import ctypes
import os
import re

import numpy as np
import pytest

import slamplay_amd as S
from slamplay_amd import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ladybug49_known_answers(ladybug_path):
    p = S.BALProblem(ladybug_path)
    assert (p.num_cameras(), p.num_points(), p.num_observations()) == (49, 7776, 31843)
    assert p.num_parameters() == 23769 == 9 * 49 + 3 * 7776
    assert p.camera_block_size() == 9 and p.point_block_size() == 3
    assert p.camera_index()[0] == 0 and p.point_index()[0] == 0
    np.testing.assert_array_equal(p.observations()[:2], [-332.65, 262.09])
    with open(ladybug_path) as fh:
        lines = fh.read().splitlines()
    assert len(lines) == 55613 == 1 + 31843 + 23769
    assert float(lines[1 + 31843]) == p.mutable_cameras()[0]
    assert float(lines[-1]) == p.mutable_points()[-1]
    # observations are sorted by point (BAL convention); tracks 2..29 long
    assert np.all(np.diff(p.point_index()) >= 0)
    counts = np.bincount(p.point_index())
    assert counts.min() == 2 and counts.max() == 29


def test_bal_round_trip(ladybug_path, tmp_path):
    p = S.BALProblem(ladybug_path)
    out = tmp_path / "rt.txt"
    p.WriteToFile(out)
    q = S.BALProblem(out)
    np.testing.assert_array_equal(q.camera_index(), p.camera_index())
    np.testing.assert_array_equal(q.point_index(), p.point_index())
    # parameters are written with %.16g as in the reference (BALProblem.cpp:169,177): 16 significant digits, so 5e-16 relative
    np.testing.assert_allclose(q.mutable_cameras(), p.mutable_cameras(), rtol=1e-15, atol=0)
    np.testing.assert_allclose(q.mutable_points(), p.mutable_points(), rtol=1e-15, atol=0)
    np.testing.assert_allclose(q.observations(), p.observations(), rtol=1e-5)  # %g, as the reference
    legacy = tmp_path / "legacy.txt"
    p.WriteToFile(legacy, legacy_header=True)
    with open(legacy) as fh:
        assert fh.readline().split() == ["49", "49", "7776", "31843"]


def test_quaternion_mode_round_trip(ladybug_path, tmp_path):
    p = S.BALProblem(ladybug_path)
    q = S.BALProblem(ladybug_path, use_quaternions=True)
    assert q.camera_block_size() == 10 and q.num_parameters() == 10 * 49 + 3 * 7776
    out = tmp_path / "q.txt"
    q.WriteToFile(out)  # written back in angle-axis form
    r = S.BALProblem(out)
    np.testing.assert_allclose(r.mutable_cameras(), p.mutable_cameras(), rtol=1e-12, atol=1e-14)


def test_ply_output(ladybug_path, tmp_path):
    p = S.BALProblem(ladybug_path)
    out = tmp_path / "x.ply"
    p.WriteToPLYFile(out)
    with open(out) as fh:
        lines = fh.read().splitlines()
    assert lines[0] == "ply" and lines[2] == "element vertex %d" % (49 + 7776) and lines[9] == "end_header"
    assert len(lines) == 10 + 49 + 7776
    assert lines[10].endswith(" 0 255 0") and lines[-1].endswith(" 255 255 255")


def test_normalize_matches_oracle(oracle_mod, ladybug_path):
    p = S.BALProblem(ladybug_path)
    cams, pts = oracle_mod.normalize(49, 7776, p.mutable_cameras(), p.mutable_points())
    p.Normalize()
    np.testing.assert_array_equal(p.mutable_cameras(), cams)
    np.testing.assert_array_equal(p.mutable_points(), pts)


def test_perturb_matches_oracle(oracle_mod, ladybug_path):
    p = S.BALProblem(ladybug_path)
    p.Normalize()
    cams, pts = oracle_mod.perturb(49, 7776, p.mutable_cameras(), p.mutable_points(), 0.1, 0.5, 0.5, seed=1)
    oracle_mod.lib().oracle_srand(1)  # glibc's default seed, as the reference never calls srand
    p.Perturb(0.1, 0.5, 0.5)
    np.testing.assert_array_equal(p.mutable_cameras(), cams)
    np.testing.assert_array_equal(p.mutable_points(), pts)


def test_synthetic_configs_deterministic():
    a = S.synthetic(1, seed=5)
    b = S.synthetic(1, seed=5)
    c = S.synthetic(1, seed=6)
    np.testing.assert_array_equal(a.observations(), b.observations())
    np.testing.assert_array_equal(a.mutable_cameras(), b.mutable_cameras())
    assert not np.array_equal(a.observations(), c.observations())
    assert a.num_cameras() == 16 and 1800 < a.num_points() <= 2000 and 13000 < a.num_observations() < 17000


def test_synthetic_noise_level(oracle_mod):
    for cfg in (1, 2, 4):
        p, gt = S.synthetic(cfg, seed=2, ground_truth=True,
                            **({"num_points": 3000} if cfg == 2 else {}))
        nc = p.num_cameras()
        r, _, _ = oracle_mod.eval_residual_jacobian(p.camera_index(), p.point_index(), p.observations(),
                                                   gt[:9 * nc], gt[9 * nc:])
        assert 0.45 < r.std() < 0.55  # pixel noise sigma 0.5
        # every observation is in front of its camera (Snavely: p_z < 0)
        cams = gt[:9 * nc].reshape(nc, 9)
        pts = gt[9 * nc:].reshape(-1, 3)
        for k in range(0, p.num_observations(), 97):
            c, q = p.camera_index()[k], p.point_index()[k]
            pz = (oracle_mod.rotate_point(cams[c, :3], pts[q]) + cams[c, 3:6])[2]
            assert pz < 0
        assert np.bincount(p.point_index()).min() >= 2


def test_large_config_sizes():
    p = S.synthetic(3, seed=1)
    assert p.num_cameras() == 1500 and p.num_observations() == 1_200_000
    counts = np.bincount(p.point_index())
    assert counts.min() >= 2 and counts.max() <= 60


def _header_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include", "mba")):
        with open(os.path.join(ROOT, "include", "mba", fn)) as fh:
            text = fh.read()
        if fn.endswith(".h") and 'extern "C"' in text:
            names |= set(re.findall(r"^\s*(?:int|int32_t|void|const char\*)\s+(ba_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    names = _header_functions()
    assert len(names) > 20
    L = C.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert L.ba_abi_version() == 2


def test_full_report_text():
    """ba_summary_full_report lays out a summary as Ceres's Summary::FullReport() (the reference
    driver prints it after ceres::Solve): size table from the problem, costs, step counts, the
    trace's device time and the termination test with its value and tolerance."""
    s, rows = S._summary(4)
    s.initial_cost, s.final_cost = 4.185660e+06, 1.803390e+04
    s.num_iterations, s.num_successful_steps, s.num_unsuccessful_steps = 3, 2, 0
    s.termination_type, s.linear_solver_used, s.total_time_ms = 0, 2, 250.0
    s.termination_reason, s.termination_value = 2, 1.769717e-09
    s.total_linear_solver_iterations = 37
    rows[2].cumulative_time_in_seconds = 0.125
    o = S.default_options()
    text = S.full_report(s, 16, 22106, 83718, o)
    table = {ln[:25].strip(): ln[25:].split() for ln in text.splitlines() if len(ln) > 25}
    assert table["Parameter blocks"] == ["22122", "22122"]
    assert table["Parameters"] == [str(9 * 16 + 3 * 22106)] * 2
    assert table["Residual blocks"] == ["83718", "83718"] and table["Residuals"] == ["167436", "167436"]
    assert table["Linear solver"] == ["SPARSE_SCHUR", "ITERATIVE_SCHUR"]
    assert table["Linear solver ordering"] == ["AUTOMATIC", "22106,16"]
    assert float(table["Initial"][0]) == 4.185660e+06 and float(table["Final"][0]) == 1.803390e+04
    assert float(table["Change"][0]) == pytest.approx(4.185660e+06 - 1.803390e+04, rel=1e-6)
    assert table["Minimizer iterations"] == ["2"] and table["Successful steps"] == ["2"]
    assert table["Linear solver iterations"] == ["37"]
    assert float(table["Minimizer (device)"][0]) == 0.125 and float(table["Total"][0]) == 0.25
    assert text.rstrip().endswith("CONVERGENCE (Function tolerance reached. |cost_change|/cost: 1.769717e-09 "
                                  "<= 1.000000e-06)")
    s.termination_type, s.termination_reason, s.termination_value = 1, 1, 50
    assert "NO_CONVERGENCE (Maximum number of iterations reached. Number of iterations: 50.)" in \
        S.full_report(s, 16, 22106, 83718, o)
    # truncation: NUL-terminated prefix, full length reported
    n = C.i64(0)
    buf = ctypes.create_string_buffer(11)
    assert C.lib().ba_summary_full_report(ctypes.byref(s), 1, 1, 1, None, buf, 11, ctypes.byref(n)) == 0
    assert len(buf.value) == 10 and n.value > 100
    assert C.lib().ba_summary_full_report(None, 1, 1, 1, None, buf, 11, None) == 1


def test_invalid_arguments_rejected_before_device_use(ladybug_path):
    p = S.synthetic(1, seed=1)
    ci = p.camera_index().copy()
    ci[5] = 99
    bad = S.BALProblem.from_arrays(ci, p.point_index(), p.observations(), p.mutable_cameras(), p.mutable_points())
    with pytest.raises(S.BAError) as e:
        S.SolveBA(bad)
    assert e.value.status == 1 and "camera index 99" in str(e.value)
    obs = p.observations().copy()
    obs[7] = np.nan
    bad = S.BALProblem.from_arrays(p.camera_index(), p.point_index(), obs, p.mutable_cameras(), p.mutable_points())
    with pytest.raises(S.BAError) as e:
        S.SolveBA(bad)
    assert e.value.status == 3
    q = S.BALProblem(ladybug_path, use_quaternions=True)
    with pytest.raises(S.BAError) as e:
        S.SolveBA(q)
    assert e.value.status == 1 and "camera_block_size" in str(e.value)
    with pytest.raises(S.BAError) as e:
        S.SolveBA(p, S.default_options(loss=7))
    assert e.value.status == 1
    with pytest.raises(S.BAError) as e:
        S.BALProblem("/nonexistent/problem.txt")
    assert e.value.status == 7
    assert C.lib().ba_server_cancel(None, 1) == 1 and C.lib().ba_server_evict(None, 1) == 1


def test_no_device_is_an_error_not_a_fallback():
    if S.device_count() > 0:
        pytest.skip("a GPU is present")
    p = S.synthetic(1, seed=1)
    with pytest.raises(S.BAError) as e:
        S.SolveBA(p)
    assert e.value.status == 5


def test_g2o_mode_rejected_by_batch_and_options():
    """The batched-window and server engines run the Ceres mode only, and g2o options are checked."""
    wins = S.local_ba_windows(2, seed=3)
    with pytest.raises(C.BAError, match="Ceres mode only"):
        S.Batch(wins, S.default_options(mode=1))
    for bad in (dict(mode=2), dict(mode=1, g2o_max_trials=0), dict(mode=1, g2o_tau=0.0)):
        with pytest.raises(C.BAError):
            S.Batch(wins, S.default_options(**bad))


def test_stereo_batch_inputs_rejected():
    """ba_stereo_batch_solve checks every window before any device work and names the bad one."""
    wins = [S.stereo_window(seed=40 + i, num_keyframes=3, num_points=20) for i in range(3)]
    wins[1] = dict(wins[1], side=wins[1]["side"].copy())
    wins[1]["side"][3] = 5
    with pytest.raises(C.BAError, match="window 1.*side"):
        S.stereo_batch_solve(wins)
    assert S.stereo_batch_solve([]) == []


def test_stereo_inputs_rejected():
    """ba_stereo_solve checks the side flags and the focal lengths before any device work."""
    w = S.stereo_window(seed=4, num_keyframes=3, num_points=20)
    args = {k: w[k] for k in ("pose_index", "point_index", "side", "observations", "poses", "points", "K", "extrinsics")}
    bad = dict(args)
    bad["side"] = w["side"].copy()
    bad["side"][0] = 2
    with pytest.raises(C.BAError, match="side"):
        S.stereo_solve(**bad)
    with pytest.raises(C.BAError, match="focal"):
        S.stereo_solve(**dict(args, K=np.array([0.0, 457.0, 367.0, 248.0])))


DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "slamplay_amd", "examples",
                      "ba_driver")


def test_cpp_driver_builds_and_reports_usage(tmp_path):
    """The reference driver's main linked against the drop-in header and library: it links and
    starts; bad arguments and an unreadable file fail before any device work."""
    import subprocess
    assert os.access(DRIVER, os.X_OK), "build the example with make -C slamplay_amd/csrc"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([DRIVER, str(tmp_path / "missing.txt"), str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "missing.txt" in r.stderr
