"""Pins of the heat-conduction oracle (oracle/falcon_oracle.c, orc_heat_*) on the CPU.

The reference has no heat operator (3_6/ModHeatConduction.f90:15-16), so the
oracle is pinned by mathematics rather than by reference vectors:
  * cell volumes sum to the closed-form volume of each Yin-Yang panel;
  * for f = r the operator equals the integrated -div(kappa grad f) of every
    cell to rounding (linear fields are exact through every kind of ghost);
  * away from interpolated ghosts it converges at second order to
    -div(kappa grad f) for f = r^4 (radial fluxes) and for the Cartesian f = z
    (whose angular terms must cancel against the radial ones), kappa = kappa0
    (r/r_ref)^-3;
  * the FCG solution equals a dense direct solve of the assembled matrix;
  * the V-cycle is linear (with its coarsest-level CG run to convergence) and
    contracts the residual; its linear prolongation is exact for coarse fields
    linear in the block's index space.
"""
import numpy as np
import pytest

import pyoracle
from falcon_amd import configs
from falcon_amd.core import R_SUN

KAPPA0, RREF = 1.0, 0.7


def _setup(shape, n, refined=False):
    lay = configs.cubic_refined(shape, n) if refined else configs.cubic(shape, n)
    o = pyoracle.Oracle(lay)
    return o, pyoracle.HeatOracle(o, KAPPA0, RREF)


def _cells(o):
    """Per-block cell centres r, theta, phi as (nblocks, nk, nj, ni) arrays, plus block info."""
    n, ng = o.ni, o.ng
    R = np.zeros((o.nblocks, n, n, n)); T = np.zeros_like(R); P = np.zeros_like(R)
    info = []
    for ib in range(o.nblocks):
        xi, xj, xk = o.centres(ib)
        r, t, p = xi[ng:ng + n], xj[ng:ng + n], xk[ng:ng + n]
        R[ib] = r[None, None, :]; T[ib] = t[None, :, None]; P[ib] = p[:, None, None]
        info.append(o.block_info(ib))
    return R, T, P, info


def _volumes(H):
    # dt = 1 s is ~1e17 times below the explicit limit: A 1 = V + O(eps) * V
    return H.apply(1.0, np.ones(H.shape))


def test_volumes_sum_to_panel_volume():
    o, H = _setup((16, 32, 96), 8, refined=True)
    V = _volumes(H)
    _, _, _, info = _cells(o)
    assert np.all(V > 0)
    for yin in (1, 0):
        ids = [ib for ib, (_, fl) in enumerate(info) if fl[0] == yin]
        lo = np.min([info[ib][0][:3] for ib in ids], axis=0)
        hi = np.max([info[ib][0][3:] for ib in ids], axis=0)
        exact = (hi[0] ** 3 - lo[0] ** 3) / 3.0 * (np.cos(lo[1]) - np.cos(hi[1])) * (hi[2] - lo[2])
        assert abs(V[ids].sum() / exact - 1.0) < 1e-12


def _smooth_mask(o, info):
    """Cells whose 7-point stencil reads no interpolated ghost: not at a radial
    block face (radial spacing changes between layers of the geometric split)
    and no face ghost whose donor is in the other panel (FindNode tries the Yin
    root first, so Yang blocks in the overlap take Yin donors).  Same-level
    angular neighbours in one panel share cell centres: their ghosts are copies.
    Donors come from the ghost table of the level-0 cells (first ghost layer)."""
    n, ng = o.ni, o.ng
    mask = np.ones((o.nblocks, n, n, n), bool)
    mask[:, :, :, 0] = False
    mask[:, :, :, n - 1] = False
    yin = np.array([fl[0] for _, fl in info])
    for ib in range(o.nblocks):
        gc = o.gc_iblocks(ib)                       # [K][J][I], padded by ng
        inner = slice(ng, ng + n)
        for axis, lo, hi in ((1, ng - 1, ng + n), (0, ng - 1, ng + n)):   # theta (J), phi (K)
            for gidx, cell in ((lo, 0), (hi, n - 1)):
                face = gc[:, gidx, inner] if axis == 1 else gc[gidx, :, :][inner, inner]
                if axis == 1:
                    face = face[inner, :]           # (K, I)
                cross = (face > 0) & (yin[np.maximum(face, 1) - 1] != yin[ib])
                if axis == 1:
                    mask[ib, :, cell, :] &= ~cross
                else:
                    mask[ib, cell, :, :] &= ~cross
    return mask


@pytest.mark.parametrize("refined", [False, True])
def test_linear_radial_field_exact(refined):
    """For f = r every ghost is exact (linear inter/extrapolation) and the
    two-point flux kappa S (f_c - f_nb)/D is exact, so the operator equals the
    integrated -div(kappa grad r) = kappa0 r_ref^3 (1/r_lo - 1/r_hi)(cos t_lo -
    cos t_hi)(p_hi - p_lo) of every cell, with the insulated faces' terms
    dropped.  Covers radial faces between layers of different spacing,
    coarse/fine faces and Yin-Yang ghosts."""
    o, H = _setup((16, 32, 96), 8, refined=refined)
    R, T, P, info = _cells(o)
    n = o.ni
    rr3 = (RREF * R_SUN) ** 3
    expect = np.zeros(H.shape)
    for ib, (rng, fl) in enumerate(info):
        rf = np.linspace(rng[0], rng[3], n + 1)
        tf = np.linspace(rng[1], rng[4], n + 1)
        pf = np.linspace(rng[2], rng[5], n + 1)
        ang = (np.cos(tf[:-1]) - np.cos(tf[1:]))[None, :, None] * np.diff(pf)[:, None, None]
        lo = KAPPA0 * rr3 / rf[:-1]
        hi = KAPPA0 * rr3 / rf[1:]
        if fl[2]:
            lo = lo.copy(); lo[0] = 0.0
        if fl[1]:
            hi = hi.copy(); hi[-1] = 0.0
        expect[ib] = (lo - hi)[None, None, :] * ang
    got = H.apply(1e300, R)
    scale = np.max(np.abs(expect))
    assert np.max(np.abs(got - expect)) < 1e-9 * scale


def _divergence_error(shape, n, which):
    o, H = _setup(shape, n)
    R, T, P, info = _cells(o)
    V = _volumes(H)
    rr = RREF * R_SUN
    yin = np.array([fl[0] for _, fl in info], bool)[:, None, None, None]
    if which == "r4":
        f = R ** 4
        expect = -8.0 * KAPPA0 * rr ** 3 / R
    else:
        f = np.where(yin, R * np.cos(T), R * np.sin(T) * np.sin(P))   # Yin z = Yang y'
        expect = 3.0 * KAPPA0 * rr ** 3 * f / R ** 5
    mask = _smooth_mask(o, info)
    # dt -> infinity leaves the conduction part: sum_f c_f (f - f_nb) ~ -int div(kappa grad f)
    got = H.apply(1e300, f) / V
    return np.max(np.abs(got - expect)[mask]) / np.max(np.abs(expect)[mask])


@pytest.mark.parametrize("which", ["r4", "z"])
def test_operator_second_order(which):
    e1 = _divergence_error((16, 32, 96), 8, which)
    e2 = _divergence_error((32, 64, 192), 16, which)
    assert e1 < 1e-3 and e2 < e1 / 3.0, (e1, e2)


def test_solve_equals_dense_direct_solve():
    o, H = _setup((4, 8, 24), 4)
    N = int(np.prod(H.shape))
    dt = 100 * H.dt_explicit
    A = np.zeros((N, N))
    e = np.zeros(N)
    for c in range(N):
        e[c] = 1.0
        A[:, c] = H.apply(dt, e.reshape(H.shape)).ravel()
        e[c] = 0.0
    Tn = np.random.default_rng(3).standard_normal(H.shape)
    b = _volumes(H).ravel() / dt * Tn.ravel()
    x_direct = np.linalg.solve(A, b)
    x, it, rel = H.solve(dt, Tn, 1e-12, 200)
    assert rel <= 1e-12 and 0 < it < 200
    assert np.linalg.norm(x.ravel() - x_direct) / np.linalg.norm(x_direct) < 1e-9
    assert np.linalg.norm(b - A @ x.ravel()) / np.linalg.norm(b) < 2e-12


def test_solve_zero_rhs():
    o, H = _setup((4, 8, 24), 4)
    x, it, rel = H.solve(10 * H.dt_explicit, np.zeros(H.shape), 1e-10, 50)
    assert it == 0 and rel == 0.0 and not x.any()


def test_vcycle_linear_and_contracting():
    """The V-cycle's only nonlinear part is the coarsest level's CG (a Krylov
    method); run to convergence it is the linear coarse inverse, and the cycle
    is linear.  With the default stop (1e-2) it still contracts."""
    o, H = _setup((16, 32, 96), 8, refined=True)
    dt = 100 * H.dt_explicit
    rng = np.random.default_rng(5)
    r1, r2 = rng.standard_normal(H.shape), rng.standard_normal(H.shape)
    H.set_coarse(1e-15, 500)
    z1, z2, z12 = H.vcycle(dt, r1), H.vcycle(dt, r2), H.vcycle(dt, r1 + 2.0 * r2)
    assert np.max(np.abs(z12 - (z1 + 2.0 * z2))) <= 1e-12 * np.max(np.abs(z12))
    H.set_coarse(1e-2, 50)
    for r in (r1, r2):
        z = H.vcycle(dt, r)
        assert np.linalg.norm(r - H.apply(dt, z)) < 0.5 * np.linalg.norm(r)


def test_coarse_cg_solves_the_coarsest_level():
    """A V-cycle on a layout of 1-cell blocks is the coarsest level alone: its
    CG run to a tight stop solves the 1-cell-per-block operator (A z = r to
    1e-8; the Yin-Yang overlap makes the operator slightly unsymmetric, so CG
    stalls near there), and the default stop leaves about 1e-2."""
    o, H = _setup((4, 8, 24), 1)
    assert H.levels == 1
    dt = 10.0 * H.dt_explicit
    r = np.random.default_rng(6).standard_normal(H.shape)
    H.set_coarse(1e-15, 200)
    z = H.vcycle(dt, r)
    assert np.linalg.norm(r - H.apply(dt, z)) <= 1e-8 * np.linalg.norm(r)
    H.set_coarse(1e-2, 50)
    z = H.vcycle(dt, r)
    assert 1e-4 < np.linalg.norm(r - H.apply(dt, z)) / np.linalg.norm(r) < 2e-2


def test_linear_prolongation_exact_on_linear_fields():
    """zf + P zc with zc linear in (theta, phi) at the coarse cell centres gives
    the same linear function at the fine centres, wherever the three coarse
    face neighbours a fine cell reads are in its own panel's same-level blocks
    (Yin blocks away from the panel's theta/phi edges; theta/phi spacing is
    uniform in a panel).  Checked between every pair of levels."""
    o, H = _setup((16, 32, 96), 8)
    R, T, P, info = _cells(o)
    yin = np.array([fl[0] for _, fl in info], bool)
    lo = np.min([rng[:3] for rng, _ in info], axis=0)
    hi = np.max([rng[3:] for rng, _ in info], axis=0)
    f = lambda t, p: 0.7 - 1.3 * t + 0.4 * p
    n = o.ni
    for level in range(H.levels - 1):
        m, mc = n >> level, n >> (level + 1)
        tc = np.zeros((o.nblocks, mc, mc, mc)); pc = np.zeros_like(tc)
        tf = np.zeros((o.nblocks, m, m, m)); pf = np.zeros_like(tf)
        ok = np.zeros(tf.shape, bool)
        for ib, (rng, fl) in enumerate(info):
            ct = rng[1] + (np.arange(mc) + 0.5) * (rng[4] - rng[1]) / mc
            cp = rng[2] + (np.arange(mc) + 0.5) * (rng[5] - rng[2]) / mc
            ft = rng[1] + (np.arange(m) + 0.5) * (rng[4] - rng[1]) / m
            fp = rng[2] + (np.arange(m) + 0.5) * (rng[5] - rng[2]) / m
            tc[ib] = ct[None, :, None]; pc[ib] = cp[:, None, None]
            tf[ib] = ft[None, :, None]; pf[ib] = fp[:, None, None]
            if yin[ib]:
                okj = ~(((np.arange(m) < 2) & np.isclose(rng[1], lo[1])) | ((np.arange(m) >= m - 2) & np.isclose(rng[4], hi[1])))
                okk = ~(((np.arange(m) < 2) & np.isclose(rng[2], lo[2])) | ((np.arange(m) >= m - 2) & np.isclose(rng[5], hi[2])))
                ok[ib] = okk[:, None, None] & okj[None, :, None]
        got = H.prolong(level, f(tc, pc), np.zeros(tf.shape))
        assert ok.sum() > 0
        assert np.max(np.abs(got - f(tf, pf))[ok]) < 1e-12, level
        # and it adds to the fine field
        base = np.random.default_rng(level).standard_normal(tf.shape)
        assert np.allclose(H.prolong(level, f(tc, pc), base) - base, got, rtol=0, atol=1e-12)
