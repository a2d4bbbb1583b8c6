"""GLM divergence cleaning (3_6/ModDivB.f90, DivB_option=1) on the GPU's FP64
path, through the C-ABI, against the oracle and against the flang build of the
reference (tests/golden/ref_mixed_glm.npz, written by make_golden.py from the
reference's own ModAdvance_rk4 with ModDivB_GLM): bit-exact, psi included.
"""
import hashlib
import os

import numpy as np
import pytest

import make_golden
import pyoracle
from conftest import random_inputs, rel_err
from falcon_amd import FalconError, Shell, configs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def _full(x, ib):
    """padded state of block ib with psi as the ninth variable"""
    return np.concatenate([x.get_state(ib), x.get_psi(ib)[None]])


def _setup(x, bgs, sts, psis, glm_on):
    if isinstance(x, Shell):
        x.set_gc_all()
        x.set_divb(1 if glm_on else 0)
    else:
        x.set_glm(glm_on)
    for ib in range(len(bgs)):
        x.set_background(ib, bgs[ib])
        x.set_state(ib, sts[ib])
        if glm_on:
            x.set_psi(ib, psis[ib])


@pytest.mark.skipif(not os.path.exists(os.path.join(HERE, "golden", "ref_mixed_glm.npz")),
                    reason="fixture not generated (tests/golden/make_golden.py mixed_glm)")
def test_glm_steps_match_the_reference():
    c = make_golden.CASES["mixed_glm"]
    fx = np.load(os.path.join(HERE, "golden", "ref_mixed_glm.npz"))
    o = pyoracle.Oracle(c["layout"])
    bgs, sts = make_golden.inputs(c, lambda ib: o.centres(ib)[0], o.pad, o.nblocks)
    s = Shell(c["layout"])
    _setup(s, bgs, [st[:8] for st in sts], [st[8] for st in sts], True)
    s.sendrecv_gc(0)
    s.sendrecv_gc(0)
    assert [_digest(_full(s, ib)) for ib in range(s.nblocks)] == list(fx["sha_exchange0"])
    di = 0
    for step in range(1, c["nsteps"] + 1):
        dt = s.advance_rk4(0.5)
        print(f"[glm reference] step {step} dt {dt!r}", flush=True)
        if step in c["dumps"]:
            assert dt == fx["dts"][di]
            di += 1
            got = [_full(s, ib) for ib in range(s.nblocks)]
            assert [_digest(a) for a in got] == list(fx[f"sha_step{step}"]), step
            assert np.array_equal(got[0], fx[f"full_step{step}"][0])


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("shape", [(16, 16, 0, 0), (8, 16, 5, 9)])
@pytest.mark.parametrize("name", ["mixed", "cubic8"])
def test_glm_steps_bit_exact_vs_oracle(name, shape, graph):
    print(f"[glm] layout={name} shape={shape} graph={graph}", flush=True)
    lay = configs.small_mixed() if name == "mixed" else configs.cubic((16, 32, 96), 8)
    s, o = Shell(lay), pyoracle.Oracle(lay)
    bgs, sts = random_inputs(o.nblocks, lambda ib: o.centres(ib)[0], o.pad, 5)
    rng = np.random.default_rng(17)
    psis = [1e3 * (2.0 * rng.random(tuple(o.pad[::-1])) - 1.0) for _ in range(o.nblocks)]
    _setup(s, bgs, sts, psis, True)
    _setup(o, bgs, sts, psis, True)
    s.set_tiles(0, *shape)
    s.sendrecv_gc(0); o.exchange(0)
    if graph:
        dts = [s.advance(4, 0.5, graph=True)]
        for _ in range(4):
            do = o.step(0.5)
        assert dts[0] == do
    else:
        for step in range(4):
            dg = s.advance_rk4(0.5)
            do = o.step(0.5)
            assert dg == do, (step, dg, do)
    a = np.stack([_full(s, ib) for ib in range(o.nblocks)])
    b = np.stack([_full(o, ib) for ib in range(o.nblocks)])
    assert np.array_equal(a, b), rel_err(a, b)


def test_glm_changes_the_field_and_is_off_by_default():
    """With psi = 0 and a field of nonzero divergence, psi grows in one step
    with GLM and B departs from the step without it (grad psi enters from the
    second stage on); with GLM off psi is refused."""
    lay = configs.small_mixed()
    o = pyoracle.Oracle(lay)
    bgs, sts = random_inputs(o.nblocks, lambda ib: o.centres(ib)[0], o.pad, 6)
    zeros = [np.zeros(tuple(o.pad[::-1])) for _ in range(o.nblocks)]
    runs = []
    for glm_on in (False, True):
        s = Shell(lay)
        _setup(s, bgs, sts, zeros, glm_on)
        s.sendrecv_gc(0)
        s.advance_rk4(0.5)
        runs.append(np.stack([s.get_state(ib) for ib in range(s.nblocks)]))
        if glm_on:
            psi = np.stack([s.get_psi(ib) for ib in range(s.nblocks)])
            assert np.abs(psi).max() > 0
        else:
            with pytest.raises(FalconError):
                s.get_psi(0)
    diff = np.abs(runs[0] - runs[1]).max(axis=(0, 2, 3, 4))
    # psi starts at 0, so grad(psi) = 0 in every stage-1 input; B changes from stage 2 on
    assert diff[4:7].max() > 0


def test_glm_runs_on_the_fp64_path_only():
    s = Shell(configs.small_mixed())
    s.set_gc_all()
    s.set_divb(1)
    with pytest.raises(FalconError):
        s.set_precision(1)
    s.set_divb(0)
    s.set_precision(1)
    with pytest.raises(FalconError):
        s.set_divb(1)


def test_nan_scan_includes_psi():
    """With GLM on, falcon_check_state scans psi as the ninth variable of
    ModCheck_primitive's order (variable innermost: psi after the eight)."""
    lay = configs.small_mixed()
    o = pyoracle.Oracle(lay)
    bgs, sts = random_inputs(o.nblocks, lambda ib: o.centres(ib)[0], o.pad, 8)
    zeros = [np.zeros(tuple(o.pad[::-1])) for _ in range(o.nblocks)]
    s = Shell(lay)
    _setup(s, bgs, sts, zeros, True)
    assert s.check_state(0) == (0, None)
    psi = np.zeros(tuple(o.pad[::-1]))
    psi[5, 4, 1] = np.nan                     # block 3, K=5, J=4, I=1: first in the scan
    s.set_psi(3, psi)
    st = s.get_state(3)
    st[7, 5, 4, 2] = np.nan                   # same K, J, next I, variable 7
    s.set_state(3, st)
    assert s.check_state(0) == (2, (3, 1, 4, 5, 8))
    psi[5, 4, 1] = 0.0
    psi[5, 4, 2] = np.nan                     # the same cell as the state's NaN: after variable 7
    s.set_psi(3, psi)
    assert s.check_state(0) == (2, (3, 2, 4, 5, 7))
