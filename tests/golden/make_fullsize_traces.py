"""Writes tests/golden/fullsize_traces.json: the FP64 CPU oracle's LM traces on the benchmark
configurations at full size (the GPU parity tests compare against these instead of re-running the
oracle for minutes on the test host).

  configuration 2: 250 cameras / 60,000 points / 473,296 observations (seed 1), explicit Schur +
                   dense Cholesky, 30 LM iterations, all termination tolerances 0;
  configuration 3: 1,500 cameras / 100,779 points / 1,200,000 observations (seed 1), implicit Schur
                   + block-Jacobi PCG (at most 50 iterations, tolerance 1e-6), 30 LM iterations;
                   and the same problem with near-exact steps (PCG to 1e-10, at most 1000
                   iterations), 10 LM iterations.

Each entry also records a SHA-256 digest of the generated problem (indices, observations and
parameters as little-endian arrays), so a change of the generator is detected before any GPU
comparison is made.  Run from the repository root: python tests/golden/make_fullsize_traces.py
(a few minutes on 8 threads); names of cases on the command line recompute only those and keep
the others."""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import slamplay_amd as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

ITERS = 30
CASES = {
    "config2_dense": dict(config=2, seed=1, options=dict(linear_solver=1)),
    "config3_pcg": dict(config=3, seed=1, options=dict(linear_solver=2, pcg_max_iterations=50, pcg_tolerance=1e-6)),
    # the same problem with near-exact linear solves: 10 LM iterations, PCG to 1e-10 (at most 1000)
    "config3_pcg_tight": dict(config=3, seed=1, iterations=10,
                              options=dict(linear_solver=2, pcg_max_iterations=1000, pcg_tolerance=1e-10)),
}


def digest(p):
    h = hashlib.sha256()
    for a in (p.camera_index(), p.point_index(), p.observations(), p.mutable_cameras(), p.mutable_points()):
        h.update(np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes())
    return h.hexdigest()


def main():
    O.set_threads(min(16, os.cpu_count() or 1))
    path = os.path.join(ROOT, "tests", "golden", "fullsize_traces.json")
    out = {}
    if len(sys.argv) > 1 and os.path.exists(path):
        with open(path) as fh:
            out = json.load(fh)
    for name, c in CASES.items():
        if len(sys.argv) > 1 and name not in sys.argv[1:]:
            continue
        iters = c.get("iterations", ITERS)
        p = S.synthetic(c["config"], seed=c["seed"])
        o = O.default_options(max_num_iterations=iters, function_tolerance=0.0, gradient_tolerance=0.0,
                              parameter_tolerance=0.0, **c["options"])
        t0 = time.time()
        r = O.solve(p.num_cameras(), p.num_points(), p.camera_index(), p.point_index(), p.observations(),
                    p.mutable_cameras(), p.mutable_points(), o)
        print("%s: %.1f s, final cost %.12e" % (name, time.time() - t0, r["final_cost"]), flush=True)
        out[name] = dict(config=c["config"], seed=c["seed"], options=c["options"], lm_iterations=iters,
                         cameras=p.num_cameras(), points=p.num_points(), observations=p.num_observations(),
                         digest=digest(p), initial_cost=r["initial_cost"], final_cost=r["final_cost"],
                         trace=[{k: row[k] for k in ("iteration", "step_is_successful", "cost", "relative_decrease",
                                                     "trust_region_radius", "linear_solver_iterations")}
                                for row in r["trace"]])
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
