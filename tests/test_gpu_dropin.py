"""Drop-in path end to end on the GPU: the reference's own Fortran modules build
the tree, blocks and tables on the host, and the Fortran shims (ModAdvance,
ModCommunication of falcon_amd/fortran over ISO_C_BINDING) step the blocks in
HBM.  The driver binary (oracle/_ref/falcon_dropin_harness) is the reference
harness built with -DFALCON_DROPIN; its dumps must equal, bit for bit, those of
the same harness stepping with the reference's numerics on the CPU.
"""
import os

import numpy as np
import pytest

import pyoracle
import ref_io
from conftest import AMP
from falcon_amd import configs

pytestmark = pytest.mark.gpu

DROPIN = os.path.join(os.path.dirname(pyoracle.HARNESS), "falcon_dropin_harness")
LAYOUTS = {"mixed": configs.small_mixed(), "ref_24x6x18": configs.SHELL_REF_24x48x144}


@pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(pyoracle.HARNESS)),
                    reason="reference harness binaries not built (make -C oracle ref dropin needs the reference sources)")
@pytest.mark.parametrize("name", list(LAYOUTS))
def test_dropin_steps_equal_reference(name):
    lay = LAYOUTS[name]
    o = pyoracle.Oracle(lay)
    rng = np.random.default_rng(21)
    bgs = [ref_io.synthetic_background(o.centres(ib)[0]) for ib in range(o.nblocks)]
    sts = [ref_io.synthetic_state(rng, o.pad, AMP) for _ in range(o.nblocks)]
    print(f"[dropin] layout={name} blocks={o.nblocks}", flush=True)
    ref = ref_io.run_harness(lay, 0.5, 3, [1, 3], 2, bgs, sts)
    gpu = ref_io.run_harness(lay, 0.5, 3, [1, 3], 2, bgs, sts, exe=DROPIN)
    for b1, b2 in zip(ref["blocks"], gpu["blocks"]):
        assert np.array_equal(b1["gcib"], b2["gcib"])
    for (s1, dt1, st1), (s2, dt2, st2) in zip(ref["dumps"], gpu["dumps"]):
        assert s1 == s2 and dt1 == dt2, (s1, dt1, dt2)
        for ib, (a, b) in enumerate(zip(st1, st2)):
            assert np.array_equal(a, b), (name, s1, ib)


@pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(pyoracle.HARNESS)),
                    reason="reference harness binaries not built (make -C oracle ref dropin needs the reference sources)")
def test_dropin_glm_steps_equal_reference():
    """DivB_option=1: the reference's ModAdvance_rk4 calls ModDivB_GLM; the
    shims push psi (primitive_IV(:,:,:,psi_)) to the device with the state and
    step it there.  States and psi equal the reference's after 1 and 3 steps."""
    lay = LAYOUTS["mixed"]
    o = pyoracle.Oracle(lay)
    rng = np.random.default_rng(22)
    bgs = [ref_io.synthetic_background(o.centres(ib)[0]) for ib in range(o.nblocks)]
    sts = [ref_io.synthetic_state(rng, o.pad, list(AMP) + [1e3]) for _ in range(o.nblocks)]
    print(f"[dropin glm] blocks={o.nblocks}", flush=True)
    ref = ref_io.run_harness(lay, 0.5, 3, [1, 3], 2, bgs, sts)
    gpu = ref_io.run_harness(lay, 0.5, 3, [1, 3], 2, bgs, sts, exe=DROPIN)
    for (s1, dt1, st1), (s2, dt2, st2) in zip(ref["dumps"], gpu["dumps"]):
        assert s1 == s2 and dt1 == dt2, (s1, dt1, dt2)
        for ib, (a, b) in enumerate(zip(st1, st2)):
            assert a.shape[0] == 9 and np.array_equal(a, b), (s1, ib)
