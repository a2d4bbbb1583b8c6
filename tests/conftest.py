import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfalcon_hip.so)")


def _ensure_built():
    from falcon_amd import core
    if not os.path.exists(core.LIB_PATH):
        from falcon_amd import build
        build.build()
    import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()


_ensure_built()


def gpu_available():
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    return os.path.exists("/dev/kfd")


AMP = np.array([1e-3, 1e2, 1e2, 1e2, 1e1, 1e1, 1e1, 1e4])


def random_inputs(nblocks, centres, pad, seed=1):
    """Seeded background and padded state for each block (test inputs)."""
    import ref_io
    rng = np.random.default_rng(seed)
    bgs, sts = [], []
    for ib in range(nblocks):
        xi = centres(ib)
        bgs.append(ref_io.synthetic_background(xi))
        sts.append(ref_io.synthetic_state(rng, pad, AMP))
    return bgs, sts


def rel_err(a, b):
    """max|a-b| / max|b| per variable (axis 1 = variable for stacked blocks)."""
    a = np.asarray(a); b = np.asarray(b)
    out = []
    for v in range(8):
        av, bv = a[:, v], b[:, v]
        m = np.max(np.abs(bv))
        out.append(float(np.max(np.abs(av - bv)) / m) if m > 0 else float(np.max(np.abs(av - bv))))
    return np.array(out)
