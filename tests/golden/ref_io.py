"""Input/output of oracle/_ref/falcon_ref_harness (the flang build of the reference).

Used by tests/golden/make_golden.py (fixture generation, in the build container)
and by CPU tests that re-run the harness when it is present.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "falcon_ref_harness")

R_SUN = float(np.float32(6.955e10))


def synthetic_background(xi, seed_shift=0.0):
    """n=1.5 polytrope profiles at radii xi (cm): rows g, rho0, p0, T0, gamma1, gamma3, heat, xi.

    Test input only.  Built from +,-,*,/ and sqrt alone (all correctly rounded),
    so the same xi gives the same bits on every machine and numpy version."""
    G, M = 6.674e-8, 1.989e33
    x = xi / R_SUN
    beta = 14.86
    zeta = 1.0 + beta * (1.0 / x - 1.0)
    sz = np.sqrt(zeta)
    rho_b = 1.0e-2
    pb_over_rhob = G * M / (2.5 * beta * R_SUN)
    rho0 = rho_b * (zeta * sz)                 # zeta**1.5
    p0 = rho_b * pb_over_rhob * (zeta * zeta * sz)   # zeta**2.5
    T0 = 1.0e4 * zeta
    g = G * M / (xi * xi)
    gam = np.full_like(xi, 5.0 / 3.0)
    heat = 1.0e-3 * (x - 0.85 + seed_shift)
    xirs = 1.0 + 0.5 * x
    return np.stack([g, rho0, p0, T0, gam, gam * 1.01, heat, xirs])


def synthetic_state(rng, pad, amp):
    """Random padded state (len(amp), NK, NJ, NI), every variable nonzero;
    uniform draws scaled per variable (uniform doubles are stable across numpy
    versions)."""
    s = 2.0 * rng.random((len(amp),) + tuple(pad[::-1])) - 1.0
    return s * np.asarray(amp, dtype=np.float64)[:, None, None, None]


def write_input(path, layout, cfl, nsteps, dumps, mode, blocks_bg=None, blocks_state=None, nvar=8):
    """nvar 9: psi is the ninth variable and the harness runs with GLM
    divergence cleaning (DivB_option=1)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", layout["ni"], layout["nj"], layout["nk"], layout["ng"], nvar))
        f.write(struct.pack("<2d", layout["r_rsun"][0] * R_SUN, layout["r_rsun"][1] * R_SUN))
        amrs = [op for op in layout.get("ops", []) if op[0] == "amr"]
        if len(amrs) != len(layout.get("ops", [])):
            raise ValueError("the reference tree only supports binary AMR divisions")
        f.write(struct.pack("<i", len(amrs)))
        for _, flags, glob, rng6 in amrs:
            f.write(struct.pack("<4i", *[int(x) for x in flags], int(glob)))
            f.write(struct.pack("<6d", *rng6))
        f.write(struct.pack("<d", cfl))
        f.write(struct.pack("<i", nsteps))
        f.write(struct.pack("<i", len(dumps)))
        if dumps:
            f.write(struct.pack("<%di" % len(dumps), *dumps))
        f.write(struct.pack("<i", mode))
        if mode >= 2:
            for bg, st in zip(blocks_bg, blocks_state):
                # harness reads bg(-ng+1:ni+ng, 8) then st(NI,NJ,NK,nvar): C order (8,NI) / (nvar,NK,NJ,NI)
                f.write(np.ascontiguousarray(bg, dtype="<f8").tobytes())
                f.write(np.ascontiguousarray(st, dtype="<f8").tobytes())


class _Reader:
    def __init__(self, data):
        self.d = data
        self.o = 0

    def i(self, n=1):
        v = np.frombuffer(self.d, dtype="<i4", count=n, offset=self.o)
        self.o += 4 * n
        return v

    def f(self, n=1):
        v = np.frombuffer(self.d, dtype="<f8", count=n, offset=self.o)
        self.o += 8 * n
        return v


def read_output(path, layout, mode, ndumps, nvar=8):
    ni, nj, nk, ng = layout["ni"], layout["nj"], layout["nk"], layout["ng"]
    NI, NJ, NK = ni + 2 * ng, nj + 2 * ng, nk + 2 * ng
    r = _Reader(open(path, "rb").read())
    nb, nyin = (int(x) for x in r.i(2))
    out = {"nblocks": nb, "nleaf_yin": nyin, "blocks": []}
    for _ in range(nb):
        rng = r.f(6)
        flags = r.i(3)
        xi, xj, xk = r.f(NI), r.f(NJ), r.f(NK)
        out["blocks"].append({"range": rng.copy(), "flags": flags.copy(), "xi": xi.copy(), "xj": xj.copy(), "xk": xk.copy()})
    for b in out["blocks"]:
        b["gcib"] = r.i(NI * NJ * NK).reshape(NK, NJ, NI).copy()
    for b in out["blocks"]:
        nt = int(r.i()[0])
        tg = []
        for _ in range(nt):
            ib, n = (int(x) for x in r.i(2))
            ijk = r.i(3 * n).reshape(3, n).copy()
            x = r.f(3 * n).reshape(3, n).copy()
            tg.append((ib, ijk, x))
        b["targets"] = tg
    for b in out["blocks"]:
        ns = int(r.i()[0])
        sc = []
        for _ in range(ns):
            ib, n = (int(x) for x in r.i(2))
            ijk = r.i(3 * n).reshape(3, n).copy()
            x = r.f(3 * n).reshape(3, n).copy()
            sc.append((ib, ijk, x))
        b["sources"] = sc
    if mode >= 2:
        dumps = []

        def read_states():
            return [r.f(nvar * NI * NJ * NK).reshape(nvar, NK, NJ, NI).copy() for _ in range(nb)]

        step = int(r.i()[0]); dt = float(r.f()[0])
        dumps.append((step, dt, read_states()))
        out["rhs"] = [r.f(8 * ni * nj * nk).reshape(8, nk, nj, ni).copy() for _ in range(nb)]
        out["diff"] = [r.f(8 * ni * nj * nk).reshape(8, nk, nj, ni).copy() for _ in range(nb)]
        for _ in range(ndumps):
            step = int(r.i()[0]); dt = float(r.f()[0])
            dumps.append((step, dt, read_states()))
        out["dumps"] = dumps
    return out


def raise_stack_limit():
    """The reference keeps whole-block temporaries on the stack (several MB per
    array for 64x16x48 blocks): raise this process's soft stack limit to its
    hard limit, which the harness processes it starts inherit."""
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    if soft != hard:
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run_harness(layout, cfl=0.5, nsteps=0, dumps=(), mode=1, blocks_bg=None, blocks_state=None, exe=None,
                launcher=()):
    """Run the harness (launcher: e.g. ["mpirun", "-n", "2"]) and read its output."""
    raise_stack_limit()
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        nvar = len(blocks_state[0]) if blocks_state else 8
        write_input(fi, layout, cfl, nsteps, list(dumps), mode, blocks_bg, blocks_state, nvar)
        subprocess.run(list(launcher) + [exe or HARNESS, fi, fo], check=True, stdout=subprocess.DEVNULL)
        return read_output(fo, layout, mode, len(dumps), nvar)
