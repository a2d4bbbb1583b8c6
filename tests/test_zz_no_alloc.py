"""GPU test of falcon_compute_dt_noalloc, the dt entry point that allocates no
device memory (its partial minima live in the library's static scratch array),
against falcon_compute_dt, which uses the handle's buffers.  dt is a minimum, so
the bar is equality; the oracle's dt is checked too."""
import numpy as np
import pytest

import pyoracle
from falcon_amd import FalconError, Shell, configs

pytestmark = pytest.mark.gpu


def _shell(layout, seed):
    s = Shell(layout)
    s.background_polytrope(20.0)
    s.set_gc_all()
    s.init_synthetic(seed=seed)
    return s


def _oracle_dt(s, layout, cfl):
    o = pyoracle.Oracle(layout)
    for ib in range(o.nblocks):
        o.set_background(ib, s.get_background(ib))
        o.set_state(ib, s.get_state(ib))
    return o.dt(cfl)


def test_noalloc_dt_equals_compute_dt_on_two_calls():
    print("[noalloc] two shells, the second smaller", flush=True)
    lay_big, lay_small = configs.cubic((32, 64, 192), 8), configs.cubic((16, 32, 96), 16)
    big, small = _shell(lay_big, 5), _shell(lay_small, 6)
    assert big.nblocks == 1536 and small.nblocks == 24
    d1 = big.compute_dt(0.5)
    n1 = big.compute_dt_noalloc(0.5)
    assert n1 == d1 == _oracle_dt(big, lay_big, 0.5)
    # the second call has 24 partial minima where the first left 1536; the
    # first shell's finer cells give the smaller dt, so a stale partial read
    # by the second call would show
    d2 = small.compute_dt(0.5)
    n2 = small.compute_dt_noalloc(0.5)
    assert d1 < d2
    assert n2 == d2 == _oracle_dt(small, lay_small, 0.5)
    # again after steps of both precisions and with the 8-edge tiling
    big.advance(3, 0.5, graph=True)
    assert big.compute_dt_noalloc(0.5) == big.compute_dt(0.5)
    small.set_precision("fp32")
    small.set_tiles(1, 8, 8)
    small.advance(2, 0.5, graph=False)
    assert small.compute_dt_noalloc(0.5) == small.compute_dt(0.5)
    big.close()
    small.close()


def test_noalloc_dt_errors():
    print("[noalloc] errors", flush=True)
    s = Shell(configs.cubic((16, 32, 96), 16))
    s.background_polytrope(20.0)
    s.set_gc_all()
    with pytest.raises(FalconError, match="not uploaded"):
        s.compute_dt_noalloc(0.5)
    s.close()
    # 1152 blocks of 2x32x32: 4608 partials at tile edge 16, 18432 at edge 8
    lay = {"ni": 2, "nj": 32, "nk": 32, "ng": 2, "r_rsun": (0.7, 1.0), "ops": [("split", (8, 4, 18))]}
    s = _shell(lay, 7)
    assert s.nblocks == 1152
    assert s.compute_dt_noalloc(0.5) == s.compute_dt(0.5)
    s.set_tiles(0, 16, 8)
    ref = s.compute_dt(0.5)
    with pytest.raises(FalconError, match="exceed its limit"):
        s.compute_dt_noalloc(0.5)
    s.set_tiles(0, 16, 16)
    assert s.compute_dt_noalloc(0.5) == ref
    s.close()
