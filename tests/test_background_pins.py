"""Pins of the n = 1.5 polytrope background (CPU, through the C-ABI).

The configured shells use a spherical polytrope after the plane-parallel one of
1_5/ModSSM.f90:16-24 (rho0 ~ zeta^n, p0 ~ zeta^(n+1), T0 ~ zeta), with
zeta(r) = 1 + beta (r_top / r - 1) chosen for hydrostatic balance under
g = G M_sun / r^2 and a bottom-to-top density contrast.  The checks below use
what the physics fixes, not the generator's formulas: hydrostatic balance by a
finite difference of p0 across the cell centres, the affine law of rho0^(1/n)
in 1/r extrapolated to the shell's ends, the ideal-gas law, and the constants.
A dropped (n+1), a wrong beta, a sign or a swapped row fails one of them.
"""
import numpy as np
import pytest

from falcon_amd import Shell, configs

G_MSUN = 6.674e-8 * 1.989e33
KB_OVER_MP = 1.380649e-16 / 1.67262192369e-24


def _profile(contrast, rho_top=1.0e-2):
    """(r, g, rho0, p0, T0, gamma1, gamma3, heat, xi) over the padded radial
    centres of one column of blocks (every block of a radial layer carries the
    same rows), sorted by r."""
    s = Shell(configs.cubic((64, 16, 48), 16))
    s.background_polytrope(contrast, rho_top)
    rows = {}
    for ib in range(s.nblocks):
        _, _, (xi, _, _) = s.block_info(ib)
        bg = s.get_background(ib)
        for I in range(xi.size):
            rows.setdefault(float(xi[I]), bg[:, I])
    s.close()
    r = np.array(sorted(rows))
    return r, np.array([rows[x] for x in r]).T


def _layers(contrast):
    """(xi, bg) of one Yin block per radial layer, in radial order."""
    s = Shell(configs.cubic((64, 16, 48), 16))
    s.background_polytrope(contrast)
    layers = {}
    for ib in range(s.nblocks):
        rng, fl, (xi, _, _) = s.block_info(ib)
        if fl[0] == 1 and rng[0] not in layers:
            layers[rng[0]] = (xi, s.get_background(ib))
    s.close()
    assert len(layers) == 4
    return [layers[k] for k in sorted(layers)]


def test_hydrostatic_balance_to_truncation_error():
    """dp0/dr = -rho0 g: a 4th-order central difference of p0 over each block's
    uniform radial centres matches -rho0 g to its truncation error (relative
    7e-7 at most, in the bottom layer, at 16 cells per layer); a g off by
    n/(n+1) would miss by 0.4."""
    worst = 0.0
    for xi, bg in _layers(20.0):
        h = xi[1] - xi[0]
        assert np.allclose(np.diff(xi), h, rtol=1e-9, atol=0)
        p0, rho0, g = bg[2], bg[1], bg[0]
        dp = (-p0[4:] + 8 * p0[3:-1] - 8 * p0[1:-3] + p0[:-4]) / (12 * h)
        bal = -rho0[2:-2] * g[2:-2]
        worst = max(worst, float(np.max(np.abs(dp - bal) / np.abs(bal))))
    assert worst < 2e-6, worst


@pytest.mark.parametrize("contrast", [20.0, 5.0])
def test_density_contrast_and_polytropic_law(contrast):
    """rho0^(1/n) (n = 1.5) is affine in 1/r; extrapolated to the shell's ends
    it gives rho0(r_bottom) / rho0(r_top) = the contrast, and rho0(r_top) =
    rho_top."""
    r, bg = _profile(contrast)
    rho0 = bg[1]
    y = rho0 ** (1.0 / 1.5)
    x = 1.0 / r
    A = np.vstack([x, np.ones_like(x)]).T
    coef, *_ = np.linalg.lstsq(A, y, rcond=None)
    resid = y - A @ coef
    assert np.max(np.abs(resid)) < 1e-12 * np.max(np.abs(y))
    R_SUN = float(np.float32(6.955e10))
    r_lo, r_hi = 0.7 * R_SUN, 1.0 * R_SUN
    rho_lo = (coef[0] / r_lo + coef[1]) ** 1.5
    rho_hi = (coef[0] / r_hi + coef[1]) ** 1.5
    assert rho_lo / rho_hi == pytest.approx(contrast, rel=1e-10)
    assert rho_hi == pytest.approx(1.0e-2, rel=1e-10)


def test_gravity_gas_law_and_constants():
    """g r^2 = G M_sun; p0 / (rho0 T0) = k_B / (mu m_p) with mu = 0.6 at every
    radius; p0 / rho0^(1 + 1/n) constant (polytropic); gamma1 = gamma3 = 5/3,
    heat = 0, xi = 1."""
    r, bg = _profile(20.0)
    g, rho0, p0, T0, g1, g3, heat, xi = bg
    np.testing.assert_allclose(g * r * r, G_MSUN, rtol=1e-13)
    np.testing.assert_allclose(p0 / (rho0 * T0), KB_OVER_MP / 0.6, rtol=1e-13)
    K = p0 / rho0 ** (1.0 + 1.0 / 1.5)
    np.testing.assert_allclose(K, K[0], rtol=1e-12)
    assert np.all(g1 == 5.0 / 3.0) and np.all(g3 == 5.0 / 3.0)
    assert np.all(heat == 0.0) and np.all(xi == 1.0)
    # density falls outward
    assert np.all(np.diff(rho0) < 0)
