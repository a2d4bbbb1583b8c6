"""Ghost tables built on the device (tables.hip, falcon_set_device_tables)
against the host build (Layout::build, itself pinned against the oracle and
the reference's fixtures): donors (GC_iBlocks), targets in first-appearance
order and every entry's base cell, weights and rotation, bit for bit, on
uniform, locally refined and randomly refined trees, and FP64 steps on
device-built tables equal to steps on host-built ones."""
import time

import numpy as np
import pytest

from conftest import random_inputs
from falcon_amd import Shell, configs

pytestmark = pytest.mark.gpu

LAYOUTS = {
    "mixed": configs.small_mixed(),
    "ref_24x6x18": configs.SHELL_REF_24x48x144,
    "cubic8": configs.cubic((16, 32, 96), 8),
    "refined8": configs.cubic_refined((16, 32, 96), 8),
    "random_amr8": configs.cubic_random_amr((16, 32, 96), 8, seed=3),
    "amr_32x64x192_b16": configs.SHELL_AMR_32x64x192_B16,
}


def _built(lay, device):
    s = Shell(lay)
    if device:
        s.set_device_tables(True)
    t0 = time.perf_counter()
    s.set_gc_all()
    return s, time.perf_counter() - t0


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_device_tables_equal_host_tables(name):
    lay = LAYOUTS[name]
    h, th = _built(lay, False)
    d, td = _built(lay, True)
    print(f"[tables] {name} blocks={h.nblocks} host {th:.3f} s device {td:.3f} s", flush=True)
    assert d.nblocks == h.nblocks
    eh, ed = h.gc_entries(), d.gc_entries()
    for k in eh:
        assert np.array_equal(eh[k], ed[k]), (name, k)
    for ib in range(h.nblocks):
        assert np.array_equal(h.gc_iblocks(ib), d.gc_iblocks(ib)), (name, ib)
        th_, td_ = h.gc_targets(ib), d.gc_targets(ib)
        assert len(th_) == len(td_)
        for (a, x), (b, y) in zip(th_, td_):
            assert a == b and np.array_equal(x, y), (name, ib)


def test_steps_on_device_tables_equal_steps_on_host_tables():
    lay = LAYOUTS["random_amr8"]
    runs = []
    for device in (False, True):
        s, _ = _built(lay, device)
        bgs, sts = random_inputs(s.nblocks, lambda ib: s.block_info(ib)[2][0], s.pad, 9)
        for ib in range(s.nblocks):
            s.set_background(ib, bgs[ib])
            s.set_state(ib, sts[ib])
        s.sendrecv_gc(0)
        s.advance(3, 0.5, graph=False)
        runs.append(np.stack([s.get_state(ib) for ib in range(s.nblocks)]))
    assert np.array_equal(runs[0], runs[1])


def test_device_tables_on_the_50m_cell_shell():
    """The 128x256x768-per-panel shell (12288 blocks): entries equal the host
    build's (GC_iBlocks is not kept above 4096 blocks); build times printed."""
    lay = configs.SHELL_128x256x768_B16
    h, th = _built(lay, False)
    d, td = _built(lay, True)
    print(f"[tables] 50M-cell shell blocks={h.nblocks} host {th:.3f} s device {td:.3f} s", flush=True)
    eh, ed = h.gc_entries(), d.gc_entries()
    for k in eh:
        assert np.array_equal(eh[k], ed[k]), k
