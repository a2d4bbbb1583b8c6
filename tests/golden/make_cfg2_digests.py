"""Writes tests/golden/cfg2_digests.json: per-block SHA-256 digests of the FP64
state on the 64x128x384-per-panel shell after 1, 100 and 1000 steps.

The shell is the reference-expressible form of that workload: blocks of
64x16x48 cells, three global theta/phi levels, 128 blocks (the reference's
tree splits theta and phi in two per level).  Inputs are regenerated from the
seed by `inputs()` (also used by the GPU test): the polytrope rows of
ref_io.synthetic_background and a seeded white-noise state.

Digests come from the C oracle (OpenMP), which is bit-exact with the flang
build of the reference step for step; steps 1 and 100 are also run through
the reference harness (oracle/_ref/falcon_ref_harness, one process) and must
give the same digests, so the step-1000 digests rest on an oracle pinned on
this very configuration.  Run in the build container:

    python tests/golden/make_cfg2_digests.py [--threads 8] [--no-harness]
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import ref_io  # noqa: E402

ALL_R = (0.6, 0.0, 0.0, 1.1, 0.0, 0.0)
LAYOUT = {"ni": 64, "nj": 16, "nk": 48, "ng": 2, "r_rsun": (0.7, 1.0),
          "ops": [("amr", (0, 1, 1), 1, ALL_R)] * 3}
SEED = 303
# per-variable amplitudes of the white-noise state: rho1, v (cm/s), B (G), s1
AMP = [1e-7, 1e2, 1e2, 1e2, 1e1, 1e1, 1e1, 1e4]
STEPS = (1, 100, 1000)
CFL = 0.5
OUT = os.path.join(HERE, "cfg2_digests.json")


def inputs(centres, pad, nblocks):
    """(backgrounds, states) of every block; centres(ib) -> padded radial centres."""
    rng = np.random.default_rng(SEED)
    bgs = [ref_io.synthetic_background(centres(ib)) for ib in range(nblocks)]
    sts = [ref_io.synthetic_state(rng, pad, AMP) for _ in range(nblocks)]
    return bgs, sts


def digest(state, ng=2):
    """SHA-256 of one block's interior state, (8, nk, nj, ni) little-endian float64."""
    inner = state[:, ng:-ng, ng:-ng, ng:-ng]
    return hashlib.sha256(np.ascontiguousarray(inner, dtype="<f8").tobytes()).hexdigest()


def main():
    import pyoracle
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--no-harness", action="store_true")
    args = ap.parse_args()
    pyoracle.set_threads(args.threads)
    o = pyoracle.Oracle(LAYOUT)
    bgs, sts = inputs(lambda ib: o.centres(ib)[0], o.pad, o.nblocks)
    for ib in range(o.nblocks):
        o.set_background(ib, bgs[ib])
        o.set_state(ib, sts[ib])
    # the driver's two ghost fills before the first step (3_6/test1.f90:143-144),
    # as the reference harness makes them
    o.exchange(0)
    o.exchange(0)
    res = {"layout": "64x16x48 blocks, 3 global theta/phi levels, 128 blocks", "seed": SEED, "cfl": CFL,
           "source": "C oracle (OpenMP), bit-exact with the reference harness at steps 1 and 100",
           "digests": {}, "dt": {}}
    t0, n = time.time(), 0
    for target in STEPS:
        while n < target:
            dt = o.step(CFL)
            n += 1
            if n % 50 == 0:
                print(f"oracle step {n} dt {dt:.6e} {time.time() - t0:.0f} s", flush=True)
        res["digests"][str(n)] = [digest(o.get_state(ib)) for ib in range(o.nblocks)]
        res["dt"][str(n)] = dt
    if not args.no_harness:
        out = ref_io.run_harness(LAYOUT, cfl=CFL, nsteps=100, dumps=[1, 100], mode=2, blocks_bg=bgs,
                                 blocks_state=sts)
        for step, dt, states in out["dumps"][1:]:
            ref = [digest(s) for s in states]
            same = ref == res["digests"][str(step)]
            print(f"harness step {step}: dt {dt:.17g} vs oracle {res['dt'][str(step)]:.17g}; digests equal: {same}")
            if not same or dt != res["dt"][str(step)]:
                raise SystemExit("the oracle and the reference harness disagree")
        res["harness_checked_steps"] = [1, 100]
    with open(OUT, "w") as f:
        json.dump(res, f, indent=0)
    print("wrote", OUT, f"{time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
