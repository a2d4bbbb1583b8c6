"""Heat solve on the GPU (libfalcon_hip.so) against the oracle (orc_heat_*).

Operator and V-cycle are elementwise FP64 with the oracle's operation order:
bit-exact.  The Krylov dot products are two-pass device reductions (a different
summation order from the oracle's serial sums), so the FCG iterates agree to
rounding: iterations within one, solutions within 1e-8 relative.  At the full
64x128x384-per-panel size the check is the true residual.
"""
import numpy as np
import pytest

import pyoracle
from falcon_amd import HeatSolver, Shell, configs

pytestmark = pytest.mark.gpu

KAPPA0, RREF = 1.0, 0.7
LAYOUTS = {
    "cubic8": configs.cubic((16, 32, 96), 8),
    "refined8": configs.cubic_refined((16, 32, 96), 8),
    "random_amr8": configs.cubic_random_amr((16, 32, 96), 8, seed=3),
}
_cache = {}


def _oracle(name):
    if name not in _cache:
        o = pyoracle.Oracle(LAYOUTS[name])
        _cache[name] = (o, pyoracle.HeatOracle(o, KAPPA0, RREF))
    return _cache[name][1]


def _gpu(name, nsys):
    s = Shell(LAYOUTS[name])
    return s, HeatSolver(s, KAPPA0, RREF, nsys)


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_levels_and_explicit_dt(name):
    H = _oracle(name)
    s, hs = _gpu(name, 1)
    print(name, hs.levels, hs.dt_explicit, flush=True)
    assert hs.levels == H.levels
    assert hs.dt_explicit == pytest.approx(H.dt_explicit, rel=1e-14)


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_apply_bitexact(name):
    H = _oracle(name)
    s, hs = _gpu(name, 2)
    dt = np.array([10.0, 300.0]) * H.dt_explicit
    hs.set_dt(dt)
    x = np.random.default_rng(11).standard_normal(hs.shape)
    y = hs.apply(x)
    for k in range(2):
        assert np.array_equal(y[k], H.apply(dt[k], x[k])), (name, k)


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_vcycle_bitexact(name):
    H = _oracle(name)
    s, hs = _gpu(name, 2)
    dt = np.array([30.0, 100.0]) * H.dt_explicit
    hs.set_dt(dt)
    r = np.random.default_rng(12).standard_normal(hs.shape)
    z = hs.vcycle(r)
    for k in range(2):
        assert np.array_equal(z[k], H.vcycle(dt[k], r[k])), (name, k)


def _true_relres(H, dt, T, x):
    b = H.apply(dt, np.ones_like(T)) * T          # (V/dt) T to ~1e-14
    return np.linalg.norm(b - H.apply(dt, x)) / np.linalg.norm(b)


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_solve_matches_oracle(name):
    H = _oracle(name)
    s, hs = _gpu(name, 2)
    dt = np.array([30.0, 300.0]) * H.dt_explicit
    T = 1e-3 * np.random.default_rng(13).standard_normal(hs.shape)
    hs.set_dt(dt)
    hs.set_T(T)
    it, rel, sec = hs.solve(1e-10, 200)
    print(name, "iterations", it.tolist(), "relres", rel.tolist(), "seconds", sec, flush=True)
    for k in range(2):
        xo, ito, relo = H.solve(dt[k], T[k], 1e-10, 200)
        x = hs.get_T(k)
        assert rel[k] <= 1e-10 and abs(int(it[k]) - ito) <= 1, (it[k], ito)
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
        assert _true_relres(H, dt[k], T[k], x) < 2e-10


def test_batched_systems_equal_single_solves():
    """Each system's reductions use the same partition whatever nsys is, and a
    converged system is frozen (alpha = 0): batching changes no bit."""
    name = "refined8"
    H = _oracle(name)
    dt = np.array([10.0, 100.0, 300.0]) * H.dt_explicit
    T = 1e-3 * np.random.default_rng(14).standard_normal((3,) + H.shape)
    T[0] = 0.0                                      # b = 0: solution 0, no iterations
    s, hs = _gpu(name, 3)
    hs.set_dt(dt)
    hs.set_T(T)
    it, rel, _ = hs.solve(1e-10, 200)
    assert it[0] == 0 and rel[0] == 0.0 and not hs.get_T(0).any()
    for k in (1, 2):
        s1, h1 = _gpu(name, 1)
        h1.set_dt(dt[k])
        h1.set_T(T[k][None])
        it1, rel1, _ = h1.solve(1e-10, 200)
        assert it1[0] == it[k] and rel1[0] == rel[k]
        assert np.array_equal(h1.get_T(0), hs.get_T(k))


def test_full_size_shell_converges():
    """64x128x384 cells per panel in 16^3 blocks (6.3M cells), dt = 100 x the
    explicit limit, Gaussian T' (sigma 1e-3): FCG reaches 1e-10 and the true
    residual, recomputed with the operator, agrees."""
    s = Shell(configs.SHELL_64x128x384_B16)
    hs = HeatSolver(s, KAPPA0, RREF, 1)
    dt = 100.0 * hs.dt_explicit
    hs.set_dt(dt)
    T = 1e-3 * np.random.default_rng(15).standard_normal(hs.shape)
    hs.set_T(T)
    it, rel, sec = hs.solve(1e-10, 300)
    print("full size: cells", hs.cells, "levels", hs.levels, "iterations", int(it[0]), "relres", float(rel[0]),
          "seconds", sec, flush=True)
    assert rel[0] <= 1e-10 and it[0] < 100
    x = hs.get_T(0)[None]
    b = hs.apply(np.ones_like(T)) * T
    assert np.linalg.norm(b - hs.apply(x)) / np.linalg.norm(b) < 2e-10


def test_graph_replays_equal_eager_solves():
    """Each FCG iteration is a replayed hipGraph: repeated solves on one handle,
    with dt and T' changed between them, equal the eagerly launched solves bit
    for bit (iterations, residuals, solutions)."""
    name = "refined8"
    H = _oracle(name)
    s, hs = _gpu(name, 2)
    rng = np.random.default_rng(16)
    for rep, fac in enumerate((30.0, 300.0, 100.0)):
        dt = np.array([fac, 2.0 * fac]) * H.dt_explicit
        T = 1e-3 * rng.standard_normal(hs.shape)
        hs.set_dt(dt)
        hs.set_T(T)
        hs.set_graph(False)
        it_e, rel_e, _ = hs.solve(1e-10, 200)
        x_e = [hs.get_T(k) for k in range(2)]
        hs.set_graph(True)
        it_g, rel_g, _ = hs.solve(1e-10, 200)
        print("replay", rep, it_e.tolist(), it_g.tolist(), flush=True)
        assert np.array_equal(it_e, it_g) and np.array_equal(rel_e, rel_g)
        for k in range(2):
            assert np.array_equal(hs.get_T(k), x_e[k])


def test_graph_follows_a_tolerance_change():
    """The convergence test inside the captured iteration carries the solve's
    tolerance: a solve at 1e-6 after one at 1e-10 on the same handle stops at
    1e-6 and equals the eagerly launched solve at 1e-6."""
    name = "refined8"
    H = _oracle(name)
    s, hs = _gpu(name, 2)
    dt = np.array([30.0, 300.0]) * H.dt_explicit
    T = 1e-3 * np.random.default_rng(17).standard_normal(hs.shape)
    hs.set_dt(dt)
    hs.set_T(T)
    hs.set_graph(True)
    it10, _, _ = hs.solve(1e-10, 200)
    it6_g, rel6_g, _ = hs.solve(1e-6, 200)
    hs.set_graph(False)
    it6_e, rel6_e, _ = hs.solve(1e-6, 200)
    print("tol change", it10.tolist(), it6_g.tolist(), it6_e.tolist(), flush=True)
    assert np.array_equal(it6_g, it6_e) and np.array_equal(rel6_g, rel6_e)
    assert (rel6_g <= 1e-6).all() and (it6_g < it10).all()
