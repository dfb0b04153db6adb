"""Writes tests/golden/c3_sensitivity.json: how far the FP64 CPU oracle's own result on the benchmark
configuration 3 (1,500 cameras, 1.2M observations, PCG stopped at 50 iterations / tolerance 1e-6,
30 LM iterations, all termination tolerances 0) moves when one input parameter is moved by one ulp.

The unperturbed run is the committed trace (fullsize_traces.json, "config3_pcg"); each trial moves one
camera parameter or point coordinate (chosen by a seeded generator) to the next larger double and
records the final cost and the worst per-row cost difference.  The GPU parity test of the truncated
configuration derives its bar from this spread.  Run from the repository root (about a minute per
trial on 8 threads): python tests/golden/make_c3_sensitivity.py"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import slamplay_amd as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

TRIALS = 7


def main():
    O.set_threads(min(16, os.cpu_count() or 1))
    golden = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(golden, "fullsize_traces.json")) as fh:
        case = json.load(fh)["config3_pcg"]
    p = S.synthetic(case["config"], seed=case["seed"])
    o = O.default_options(max_num_iterations=case["lm_iterations"], function_tolerance=0.0, gradient_tolerance=0.0,
                          parameter_tolerance=0.0, **case["options"])
    base = [r["cost"] for r in case["trace"]]
    rng = np.random.default_rng(5)
    trials = []
    for t in range(TRIALS):
        cams, pts = p.mutable_cameras().copy(), p.mutable_points().copy()
        arr, what = (cams, "camera parameter") if t % 2 == 0 else (pts, "point coordinate")
        k = int(rng.integers(arr.size))
        flat = arr.reshape(-1)
        flat[k] = np.nextafter(flat[k], np.inf)
        r = O.solve(p.num_cameras(), p.num_points(), p.camera_index(), p.point_index(), p.observations(), cams, pts, o)
        rows = [abs(a["cost"] - b) / b for a, b in zip(r["trace"], base)]
        trials.append(dict(moved="%s %d" % (what, k), final_cost=r["final_cost"],
                           final_rel_diff=abs(r["final_cost"] - case["final_cost"]) / case["final_cost"],
                           worst_row_rel_diff=max(rows)))
        print(trials[-1], flush=True)
    out = dict(case="config3_pcg", base_final_cost=case["final_cost"], trials=trials,
               max_final_rel_diff=max(t["final_rel_diff"] for t in trials),
               max_row_rel_diff=max(t["worst_row_rel_diff"] for t in trials))
    with open(os.path.join(golden, "c3_sensitivity.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
