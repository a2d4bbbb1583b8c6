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
    fx = np.load(os.path.join(HERE, "golden", f"ref_{name}.npz"))
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


def test_operators_and_steps_match_reference(case):
    name, c, fx, o, bgs, sts = case
    for ib in range(o.nblocks):
        o.set_background(ib, bgs[ib])
        o.set_state(ib, sts[ib])
    o.exchange(0)
    o.exchange(0)
    fb = list(fx["full_blocks"])

    def check(label, arrays):
        assert [_digest(a) for a in arrays] == list(fx[f"sha_{label}"]), label
        for q, b in enumerate(fb):
            assert np.array_equal(arrays[b], fx[f"full_{label}"][q]), label

    check("exchange0", [o.get_state(ib) for ib in range(o.nblocks)])
    check("rhs", [o.rhs(ib) for ib in range(o.nblocks)])
    check("diff", [o.diffusion(ib) for ib in range(o.nblocks)])
    di = 0
    for s in range(1, c["nsteps"] + 1):
        dt = o.step(0.5)
        if s in c["dumps"]:
            assert dt == fx["dts"][di]
            di += 1
            check(f"step{s}", [o.get_state(ib) for ib in range(o.nblocks)])


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
