"""The oracle against the flang build of the reference (CPU).

tests/golden/ref_*.npz were written by tests/golden/make_golden.py from
oracle/_ref/falcon_ref_harness, i.e. the reference's own modules
(3_6/ModEquation, ModDiffusion, ModBoundary, ModGC, ModYinYangTree, ...)
compiled by flang.  Outputs are compared bit for bit (SHA-256 of every block).
"""
import hashlib
import os

import numpy as np
import pytest

import make_golden
import pyoracle
import ref_io

HERE = os.path.dirname(os.path.abspath(__file__))


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.fixture(scope="module", params=sorted(make_golden.CASES))
def case(request):
    name = request.param
    c = make_golden.CASES[name]
    path = os.path.join(HERE, "golden", f"ref_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"fixture not generated (tests/golden/make_golden.py {name})")
    fx = np.load(path)
    o = pyoracle.Oracle(c["layout"])
    bgs, sts = make_golden.inputs(c, lambda ib: o.centres(ib)[0], o.pad, o.nblocks)
    return name, c, fx, o, bgs, sts


def test_tables_match_reference(case):
    name, c, fx, o, _, _ = case
    assert o.nblocks == int(fx["nblocks"]) and o.nleaf_yin == int(fx["nleaf_yin"])
    for ib in range(o.nblocks):
        rng, fl = o.block_info(ib)
        assert np.array_equal(rng, fx["ranges"][ib]) and np.array_equal(fl, fx["flags"][ib])
        assert np.array_equal(o.gc_iblocks(ib), fx["gcib"][ib].astype(np.int32))
    donors, counts, ijk = fx["target_donors"], fx["target_counts"], fx["target_ijk"]
    pos, row = 0, 0
    for ib in range(o.nblocks):
        for d, t_ijk, _ in o.targets(ib):
            assert tuple(donors[row]) == (ib, d)
            n = counts[row]
            assert t_ijk.shape[1] == n
            assert np.array_equal(t_ijk, ijk[:, pos:pos + n].astype(np.int32))
            pos += n
            row += 1
    assert row == len(donors)


def _state(o, ib, glm):
    """the block's padded state, psi as a ninth variable with GLM"""
    u = o.get_state(ib)
    return np.concatenate([u, o.get_psi(ib)[None]]) if glm else u


def test_operators_and_steps_match_reference(case):
    """RHS, diffusion, exchange and steps; with GLM (mixed_glm) the steps run
    ModDivB_GLM inside ModAdvance_rk4 and psi is compared with the rest."""
    name, c, fx, o, bgs, sts = case
    glm = c.get("nvar", 8) == 9
    o.set_glm(glm)
    for ib in range(o.nblocks):
        o.set_background(ib, bgs[ib])
        o.set_state(ib, sts[ib][:8])
        if glm:
            o.set_psi(ib, sts[ib][8])
    o.exchange(0)
    o.exchange(0)
    fb = list(fx["full_blocks"])

    def check(label, arrays):
        assert [_digest(a) for a in arrays] == list(fx[f"sha_{label}"]), label
        for q, b in enumerate(fb):
            assert np.array_equal(arrays[b], fx[f"full_{label}"][q]), label

    check("exchange0", [_state(o, ib, glm) for ib in range(o.nblocks)])
    check("rhs", [o.rhs(ib) for ib in range(o.nblocks)])
    check("diff", [o.diffusion(ib) for ib in range(o.nblocks)])
    di = 0
    for s in range(1, c["nsteps"] + 1):
        dt = o.step(0.5)
        if s in c["dumps"]:
            assert dt == fx["dts"][di]
            di += 1
            check(f"step{s}", [_state(o, ib, glm) for ib in range(o.nblocks)])


@pytest.mark.skipif(not os.path.exists(ref_io.HARNESS), reason="flang build of the reference not present")
def test_live_reference_fresh_seed():
    """A fresh seed straight through the flang-built reference (build container only)."""
    c = dict(make_golden.CASES["mixed"])
    c["seed"] = 9001
    o = pyoracle.Oracle(c["layout"])
    bgs, sts = make_golden.inputs(c, lambda ib: o.centres(ib)[0], o.pad, o.nblocks)
    out = ref_io.run_harness(c["layout"], cfl=0.5, nsteps=2, dumps=[2], mode=2, blocks_bg=bgs, blocks_state=sts)
    for ib in range(o.nblocks):
        o.set_background(ib, bgs[ib])
        o.set_state(ib, sts[ib])
    o.exchange(0)
    o.exchange(0)
    o.step(0.5)
    dt = o.step(0.5)
    assert dt == out["dumps"][1][1]
    for ib in range(o.nblocks):
        assert np.array_equal(o.get_state(ib), out["dumps"][1][2][ib])


MPIRUN = "/opt/conda/bin/mpirun"


@pytest.mark.skipif(not (os.path.exists(ref_io.HARNESS) and os.path.exists(MPIRUN)),
                    reason="flang/MPICH build of the reference not present")
@pytest.mark.parametrize("ranks", [2, 3])
def test_reference_under_mpirun_matches_one_rank(ranks):
    """The reference driver with its own ModCommunication/ModAdvance under
    `mpirun -n N`: blocks split over ranks by ModAllocation_Init, remote ghosts
    sent and received over MPI.  Tables, the initial two-fill state (3_6/test1.f90:
    143-144), the RHS, the diffusion and the first step's dt equal the one-rank
    run's bit for bit.

    The states after a step are not compared: the reference's receive path
    copies each received message into the targets' lists right after posting
    MPI_IRECV (3_6/ModCommunication.f90:731-739), before MPI_WAITALL (:764), so
    a remote ghost gets the message of the previous fill.  The initial fills
    are two calls on one state, so they come out right; every fill during a
    step on more than one rank uses the previous fill's remote values.  The
    one-rank run (every donor local) and this library (one process per GPU,
    every block resident) have no remote ghosts."""
    c = make_golden.CASES["mixed"]
    o = pyoracle.Oracle(c["layout"])
    bgs, sts = make_golden.inputs(c, lambda ib: o.centres(ib)[0], o.pad, o.nblocks)
    one = ref_io.run_harness(c["layout"], cfl=0.5, nsteps=1, dumps=[1], mode=2, blocks_bg=bgs, blocks_state=sts)
    many = ref_io.run_harness(c["layout"], cfl=0.5, nsteps=1, dumps=[1], mode=2, blocks_bg=bgs, blocks_state=sts,
                              launcher=[MPIRUN, "-n", str(ranks)])
    assert one["nblocks"] == many["nblocks"]
    for a, b in zip(one["blocks"], many["blocks"]):
        assert np.array_equal(a["range"], b["range"]) and np.array_equal(a["gcib"], b["gcib"])
        assert [(t[0], t[1].tolist(), t[2].tolist()) for t in a["targets"]] == \
               [(t[0], t[1].tolist(), t[2].tolist()) for t in b["targets"]]
        # remote sources are appended in arrival order (MPI_ANY_SOURCE): same set
        key = lambda t: (t[0], t[1].tobytes(), t[2].tobytes())
        assert sorted(map(key, a["sources"])) == sorted(map(key, b["sources"]))
    for x, y in zip(one["dumps"][0][2], many["dumps"][0][2]):
        assert np.array_equal(x, y)
    for lab in ("rhs", "diff"):
        for x, y in zip(one[lab], many[lab]):
            assert np.array_equal(x, y), lab
    assert one["dumps"][1][1] == many["dumps"][1][1]
