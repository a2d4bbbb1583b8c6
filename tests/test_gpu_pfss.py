"""PFSS potential field on the GPU (libfalcon_hip.so, falcon_pfss_*) against
the oracle (orc_heat_* with the Dirichlet top, orc_pfss_field) and against the
closed-form monopole field.

The operator, the V-cycle and the field are elementwise FP64 in the oracle's
operation order: bit-exact.  The solve's dot products are device reductions in
another summation order, so solutions agree within 1e-8 relative and
iteration counts within one.  The point-charge magnetogram is checked against
a numpy restatement of the same sum.
"""
import numpy as np
import pytest

import pyoracle
from falcon_amd import HeatSolver, Pfss, Shell, configs

pytestmark = pytest.mark.gpu

LAYOUTS = {
    "single8": configs.cubic((8, 16, 48), 8),
    "cubic8": configs.cubic((16, 32, 96), 8),
    "refined8": configs.cubic_refined((16, 32, 96), 8),
    "random_amr8": configs.cubic_random_amr((16, 32, 96), 8, seed=3),
}
_cache = {}


def _oracle(name):
    if name not in _cache:
        o = pyoracle.Oracle(LAYOUTS[name])
        _cache[name] = (o, pyoracle.PfssOracle(o))
    return _cache[name]


def _dipoles(R0):
    """a centred dipole along z and an off-centre tilted one below the surface"""
    xyz = np.array([[0.0, 0.0, 0.0], [0.3 * R0, -0.2 * R0, 0.5 * R0]])
    q = np.array([1.0, -0.4])
    l = np.array([[0.0, 0.0, 0.05 * R0], [0.02 * R0, 0.03 * R0, 0.0]])
    return xyz, q, l


def _r_range(o):
    info = [o.block_info(ib) for ib in range(o.nblocks)]
    return min(r[0] for r, _ in info), max(r[3] for r, _ in info), info


def _coordconv(t, p):
    """(theta, phi) of the other Yin-Yang panel's frame: x' = -x, y' = z, z' = y"""
    st, ct, sp, cp = np.sin(t), np.cos(t), np.sin(p), np.cos(p)
    t2 = np.arccos(st * sp)
    p2 = np.copysign(np.abs(np.arccos(-(st * cp) / np.sqrt((st * st) * (cp * cp) + ct * ct))), ct)
    return t2, p2


def _point_charge_br(o, P, info, R0, n, xyz, q, l):
    """Br of the pole pairs at (R0, theta face j, phi centre k) of each bottom block"""
    br = np.zeros((len(P.bottom), n, n))
    for b, ib in enumerate(P.bottom):
        rng, fl = info[ib]
        tf = np.linspace(rng[1], rng[4], n + 1)[:n]
        pc = o.centres(ib)[2][o.ng:o.ng + n]
        T, Ph = np.meshgrid(tf, pc)              # [k][j]
        if fl[0]:
            T, Ph = _coordconv(T, Ph)
        f = R0 * np.stack([np.cos(Ph) * np.sin(T), np.sin(Ph) * np.sin(T), np.cos(T)])
        for d in range(len(q)):
            for sgn in (1.0, -1.0):
                dv = f - (xyz[d] + sgn * l[d] / 2)[:, None, None]
                rr = np.sqrt((dv * dv).sum(0))
                br[b] += sgn * q[d] * (dv * f / R0).sum(0) / rr ** 3
    return br


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_bottom_blocks_areas_and_magnetogram(name):
    o, P = _oracle(name)
    s = Shell(LAYOUTS[name])
    pf = Pfss(s)
    n = pf.n
    assert list(pf.bottom) == P.bottom
    R0, _, info = _r_range(o)
    area = pf.areas()
    for b, ib in enumerate(P.bottom):
        rng = info[ib][0]
        tf = np.linspace(rng[1], rng[4], n + 1)
        ph = np.linspace(rng[2], rng[5], n + 1)
        exact = R0 ** 2 * np.outer(ph[1:] - ph[:-1], np.cos(tf[:-1]) - np.cos(tf[1:]))
        assert np.allclose(area[b], exact, rtol=1e-11, atol=0)
    xyz, q, l = _dipoles(R0)
    br = pf.dipole_magnetogram(xyz, q, l)
    ref = _point_charge_br(o, P, info, R0, n, xyz, q, l)
    assert np.abs(br - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_apply_and_vcycle_bitexact(name):
    o, P = _oracle(name)
    s = Shell(LAYOUTS[name])
    pf = Pfss(s)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(pf.shape)
    assert np.array_equal(pf.apply(x), P.apply(float("inf"), x)), name
    r = rng.standard_normal(pf.shape)
    assert np.array_equal(pf.vcycle(r), P.vcycle(float("inf"), r)), name


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_solve_and_field_match_oracle(name):
    o, P = _oracle(name)
    s = Shell(LAYOUTS[name])
    pf = Pfss(s)
    R0, _, _ = _r_range(o)
    br = pf.dipole_magnetogram(*_dipoles(R0))
    pf.set_magnetogram(br)
    it, rel, sec = pf.solve(1e-10, 500)
    phi = pf.phi()
    phi_o, it_o, rel_o = P.solve_br(br, pf.areas(), 1e-10, 500)
    print(name, it, rel, sec, it_o, rel_o, flush=True)
    assert rel < 1e-10 and abs(it - it_o) <= 1
    assert np.abs(phi - phi_o).max() <= 1e-8 * np.abs(phi_o).max()
    # the field of the GPU's own potential, bit-exact
    B = pf.field()
    assert B.shape == (o.nblocks, 3, pf.n, pf.n, pf.n)
    assert np.array_equal(B, P.field(phi, br)), name
    # and close to the field of the oracle's potential
    Bo = P.field(phi_o, br)
    assert np.abs(B - Bo).max() <= 1e-7 * np.abs(Bo).max()


def test_monopole_converges_at_second_order_on_the_gpu():
    """Br(R0) = 1 on shells of one radial block layer: phi = R0^2 (1/r - 1/R1)"""
    errs = []
    for lay in (configs.cubic((4, 8, 24), 4), configs.cubic((8, 16, 48), 8)):
        o = pyoracle.Oracle(lay)
        R0, R1, _ = _r_range(o)
        pf = Pfss(Shell(lay))
        pf.set_magnetogram(np.ones(pf.mshape))
        it, rel, _ = pf.solve(1e-12, 500)
        assert rel < 1e-12
        phi, B = pf.phi(), pf.field()
        e = np.zeros(2)
        n, ng = pf.n, o.ng
        for ib in range(o.nblocks):
            r = o.centres(ib)[0][ng:ng + n][None, None, :]
            e[0] = max(e[0], np.abs(phi[ib] - R0 ** 2 * (1 / r - 1 / R1)).max())
            e[1] = max(e[1], np.abs(B[ib, 0] - R0 ** 2 / r ** 2).max())
        errs.append(e / np.array([R0 ** 2 * (1 / R0 - 1 / R1), 1.0]))
    order = np.log2(errs[0] / errs[1])
    print(errs, order, flush=True)
    assert order[0] > 1.9 and order[1] > 1.8


def test_heat_only_calls_refuse_the_pfss_solver():
    """The handle holds one implicit solver: after falcon_pfss_setup the heat
    solve must fail loudly rather than solve the PFSS system; a heat set-up
    afterwards replaces it."""
    s = Shell(LAYOUTS["single8"])
    Pfss(s)
    assert s.L.falcon_heat_solve(s.h, 1e-10, 10, 0, None, None, None) != 0
    hs = HeatSolver(s, 1.0, 0.7, 1)
    hs.set_dt(np.array([100.0 * hs.dt_explicit]))
    assert s.L.falcon_pfss_solve(s.h, 1e-10, 10, None, None, None) != 0
