"""Writes tests/golden/ladybug_sensitivity.json: how far the FP64 CPU oracle's own result on ladybug49
moves when one input parameter is moved by one ulp, for the two pipelines the reference runs on it:

  driver -- backend/slam_ba_simple_g2o_ceres/test_bundle_adjustment_ceres.cpp: Normalize,
            srand(1), Perturb(0.1, 0.5, 0.5), SolveBA with the default (Ceres) options;
  simple -- backend/ceres/test_ceres_bundle_adjuster_simple.cpp: the problem as loaded,
            DENSE_SCHUR, Ceres defaults.

Each trial moves one camera parameter or point coordinate (chosen by a seeded generator) of the
solver's input to the next larger double, solves again and records the final cost, its distance
from the unperturbed final cost and the first trace row whose cost differs by more than 1e-9.  The
GPU parity tests of these pipelines take their final-cost bar from this spread.  Run from the
repository root (a few seconds per trial): python tests/golden/make_ladybug_sensitivity.py"""
import gzip
import json
import os
import shutil
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import slamplay_amd as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

TRIALS = 8


def problem(case, path):
    p = S.BALProblem(path)
    if case == "driver":
        p.Normalize()
        O.lib().oracle_srand(1)
        p.Perturb(0.1, 0.5, 0.5)
    return p


def solve(p, cams, pts, case):
    o = O.default_options(linear_solver=1) if case == "simple" else O.default_options()
    return O.solve(p.num_cameras(), p.num_points(), p.camera_index(), p.point_index(), p.observations(), cams, pts, o)


def main():
    O.set_threads(min(16, os.cpu_count() or 1))
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "ladybug49.txt")
    with gzip.open(os.path.join(ROOT, "tests", "golden", "ladybug49.txt.gz"), "rb") as src, open(path, "wb") as dst:
        shutil.copyfileobj(src, dst)
    out = {}
    for case in ("driver", "simple"):
        p = problem(case, path)
        base = solve(p, p.mutable_cameras().copy(), p.mutable_points().copy(), case)
        bc = np.array([r["cost"] for r in base["trace"]])
        rng = np.random.default_rng(11)
        trials = []
        for t in range(TRIALS):
            cams, pts = p.mutable_cameras().copy(), p.mutable_points().copy()
            arr, what = (cams, "camera parameter") if t % 2 == 0 else (pts, "point coordinate")
            k = int(rng.integers(arr.size))
            flat = arr.reshape(-1)
            flat[k] = np.nextafter(flat[k], np.inf)
            r = solve(p, cams, pts, case)
            c = np.array([x["cost"] for x in r["trace"]])
            n = min(len(c), len(bc))
            far = np.nonzero(np.abs(c[:n] - bc[:n]) > 1e-9 * bc[:n])[0]
            trials.append(dict(moved="%s %d" % (what, k), final_cost=r["final_cost"], iterations=len(c),
                               termination=r["termination"],
                               final_rel_diff=abs(r["final_cost"] - base["final_cost"]) / base["final_cost"],
                               first_row_apart=int(far[0]) if far.size else None))
            print(case, trials[-1], flush=True)
        out[case] = dict(base_final_cost=base["final_cost"], base_iterations=len(bc), trials=trials,
                         max_final_rel_diff=max(t["final_rel_diff"] for t in trials))
    with open(os.path.join(ROOT, "tests", "golden", "ladybug_sensitivity.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
