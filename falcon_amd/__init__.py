"""falcon_amd: MI355X-native RK4 / ghost-fill / heat-solve hot path of Falcon.

Compute runs in libfalcon_hip.so (hand-written HIP for gfx950, built in-tree by
`python -m falcon_amd.build`); this package is the host-side driver.
"""
from .core import Cnn, FalconError, HeatSolver, Pfss, Shell, format_es16, load, R_SUN  # noqa: F401
from . import configs  # noqa: F401
