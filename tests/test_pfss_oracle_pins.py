"""Pins of the PFSS oracle (oracle/falcon_oracle.c, orc_heat_* with a
Dirichlet top and orc_pfss_field) on the CPU, against closed-form potential
fields of a shell R0 < r < R1 with phi(R1) = 0 and Br(R0) given:

  * monopole, Br(R0) = 1:  phi = R0^2 (1/r - 1/R1),  B = (R0^2/r^2, 0, 0);
  * dipole, Br(R0) = cos(Theta) (Theta the Yin panel's colatitude; in the Yang
    frame cos(Theta) = sin(theta) sin(phi)):
        phi = cos(Theta) (A r + C / r^2),  C = 1 / (1/R1^3 + 2/R0^3),  A = -C/R1^3,
    which is harmonic, vanishes on r = R1 and has -dphi/dr = cos(Theta) on R0.

The right-hand side is the exact flux of Br through every bottom face (closed
form of the face integral), so what the tests see is the operator, the
boundary treatment and the field's differences alone.  With a single radial
block layer no ghost is interpolated along r, and the monopole converges at
second order in phi and Br (a mistake in the top condition, in the face areas
or in the flux sign drops the order or leaves an O(1) error); the dipole's
error also passes through the Yin-Yang ghosts, whose interpolation limits the
order to between one and two.
"""
import numpy as np

import pyoracle
from falcon_amd import configs


def _shell(shape, n):
    o = pyoracle.Oracle(configs.cubic(shape, n))
    P = pyoracle.PfssOracle(o)
    info = [o.block_info(ib) for ib in range(o.nblocks)]
    R0 = min(rng[0] for rng, _ in info)
    R1 = max(rng[3] for rng, _ in info)
    return o, P, info, R0, R1


def _bottom_faces(P, info, n):
    """theta and phi face coordinates of every bottom block."""
    for q, ib in enumerate(P.bottom):
        rng, fl = info[ib]
        yield q, ib, fl[0], np.linspace(rng[1], rng[4], n + 1), np.linspace(rng[2], rng[5], n + 1)


def _magnetogram(P, info, n, R0, kind):
    """(br, area): face areas and the face mean of Br (flux / area)."""
    br = np.zeros((len(P.bottom), n, n))
    area = np.zeros_like(br)
    for q, ib, yin, tf, pf in _bottom_faces(P, info, n):
        # [k][j]: phi rows, theta fastest
        area[q] = R0 ** 2 * np.outer(pf[1:] - pf[:-1], np.cos(tf[:-1]) - np.cos(tf[1:]))
        if kind == "monopole":
            flux = area[q]
        elif yin:   # int cos(t) sin(t) dt dp
            flux = R0 ** 2 * np.outer(pf[1:] - pf[:-1], (np.sin(tf[1:]) ** 2 - np.sin(tf[:-1]) ** 2) / 2)
        else:       # int sin(t) sin(p) sin(t) dt dp
            st = tf / 2 - np.sin(2 * tf) / 4
            flux = R0 ** 2 * np.outer(np.cos(pf[:-1]) - np.cos(pf[1:]), st[1:] - st[:-1])
        br[q] = flux / area[q]
    return br, area


def _exact(o, info, ib, n, R0, R1, kind):
    """phi and (Br, Btheta, Bphi) at the cell centres of block ib, [k][j][i]."""
    ng = o.ng
    xi, xj, xk = o.centres(ib)
    r = xi[ng:ng + n][None, None, :]
    t = xj[ng:ng + n][None, :, None]
    p = xk[ng:ng + n][:, None, None]
    one = np.ones((n, n, n))
    if kind == "monopole":
        return R0 ** 2 * (1 / r - 1 / R1) * one, [R0 ** 2 / r ** 2 * one, 0 * one, 0 * one]
    C = 1.0 / (1.0 / R1 ** 3 + 2.0 / R0 ** 3)
    A = -C / R1 ** 3
    if info[ib][1][0]:
        c, dc_t, dc_p = np.cos(t) * one, -np.sin(t) * one, 0 * one
    else:
        c, dc_t, dc_p = np.sin(t) * np.sin(p) * one, np.cos(t) * np.sin(p) * one, np.sin(t) * np.cos(p) * one
    f = A * r + C / r ** 2
    return c * f, [-c * (A - 2 * C / r ** 3), -dc_t * f / r, -dc_p * f / (r * np.sin(t))]


def _errors(shape, n, kind):
    """max |error| of phi and of the three B components, relative to max |phi|
    and max |B| of the exact field."""
    o, P, info, R0, R1 = _shell(shape, n)
    br, area = _magnetogram(P, info, n, R0, kind)
    phi, its, rel = P.solve_br(br, area, 1e-12, 500)
    assert rel < 1e-12 and its < 100
    B = P.field(phi, br)
    e = np.zeros(4)
    top = np.zeros(2)
    for ib in range(o.nblocks):
        ex, Bex = _exact(o, info, ib, n, R0, R1, kind)
        e[0] = max(e[0], np.abs(phi[ib] - ex).max())
        top[0] = max(top[0], np.abs(ex).max())
        for a in range(3):
            e[1 + a] = max(e[1 + a], np.abs(B[ib, a] - Bex[a]).max())
            top[1] = max(top[1], np.abs(Bex[a]).max())
    return np.array([e[0] / top[0], e[1] / top[1], e[2] / top[1], e[3] / top[1]])


# one radial block layer: (nr, nt, np) cells per panel with nr = block edge
COARSE, FINE = ((4, 8, 24), 4), ((8, 16, 48), 8)


def test_monopole_converges_at_second_order():
    e1 = _errors(*COARSE, "monopole")
    e2 = _errors(*FINE, "monopole")
    order = np.log2(e1[:2] / e2[:2])
    assert e2[0] < 2e-3 and e2[1] < 3e-3
    assert order[0] > 1.9 and order[1] > 1.8, order
    # a field of r alone: the angular differences of phi cancel to rounding
    assert e1[2] < 1e-10 and e1[3] < 1e-10 and e2[2] < 1e-10 and e2[3] < 1e-10


def test_dipole_converges():
    e1 = _errors(*COARSE, "dipole")
    e2 = _errors(*FINE, "dipole")
    order = np.log2(e1 / e2)
    assert np.all(e2 < 4e-3), e2
    assert order[0] > 1.5 and order[1] > 1.8 and order[2] > 0.8 and order[3] > 0.8, order


def test_top_dirichlet_makes_the_operator_definite():
    """Without the source-surface condition the Laplacian with insulated faces
    annihilates constants; with it a constant potential is pushed back only
    through the top faces, by 2 kappa S / dr per cell (the ghost -phi_m at the
    distance dr of the cell's centre)."""
    o, P, info, R0, R1 = _shell(*COARSE)
    n = COARSE[1]
    y = P.apply(float("inf"), np.ones(P.shape))
    for ib in range(o.nblocks):
        rng, fl = info[ib]
        inner = y[ib][:, :, :n - 1] if fl[1] else y[ib]
        assert np.abs(inner).max() < 1e-6 * np.abs(y).max()
        if fl[1]:
            xi = o.centres(ib)[0][o.ng:o.ng + n]
            dr = 2 * (rng[3] - xi[-1])
            tf = np.linspace(rng[1], rng[4], n + 1)
            pf = np.linspace(rng[2], rng[5], n + 1)
            S = R1 ** 2 * np.outer(pf[1:] - pf[:-1], np.cos(tf[:-1]) - np.cos(tf[1:]))
            assert np.allclose(y[ib][:, :, n - 1], 2 * S / dr, rtol=1e-12)
