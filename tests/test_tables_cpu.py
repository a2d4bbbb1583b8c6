"""Ghost tables of the product (host C++ in libfalcon_hip.so) on the CPU.

Integer tables must be bit-exact: against the oracle, against the reference's
fixtures, and against a brute-force point location on small trees.
"""
import os

import numpy as np
import pytest

import make_golden
import pyoracle
from falcon_amd import Shell, configs

HERE = os.path.dirname(os.path.abspath(__file__))

LAYOUTS = {
    "mixed": configs.small_mixed(),
    "ref128": configs.SHELL_REF_24x48x144,
    "cubic_2x4x12": configs.cubic((16, 32, 96), 8),
    "cubic_3x6x18": configs.SHELL_24x48x144_B8,
    "random_amr": configs.cubic_random_amr((16, 32, 96), 8, seed=3),
}


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_tables_equal_oracle(name):
    lay = LAYOUTS[name]
    s, o = Shell(lay), pyoracle.Oracle(lay)
    s.build_tables()
    assert (s.nblocks, s.nleaf_yin) == (o.nblocks, o.nleaf_yin)
    for ib in range(o.nblocks):
        rng, fl, (xi, xj, xk) = s.block_info(ib)
        rng2, fl2 = o.block_info(ib)
        xi2, xj2, xk2 = o.centres(ib)
        assert np.array_equal(rng, rng2) and np.array_equal(fl, fl2)
        assert np.array_equal(xi, xi2) and np.array_equal(xj, xj2) and np.array_equal(xk, xk2)
        assert np.array_equal(s.gc_iblocks(ib), o.gc_iblocks(ib))
        t1, t2 = s.gc_targets(ib), o.targets(ib)
        assert len(t1) == len(t2)
        for (d1, ijk1), (d2, ijk2, _) in zip(t1, t2):
            assert d1 == d2 and np.array_equal(ijk1, ijk2)


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_tables_equal_reference_fixture(name):
    c = make_golden.CASES[name]
    fx = np.load(os.path.join(HERE, "golden", f"ref_{name}.npz"))
    s = Shell(c["layout"])
    s.build_tables()
    assert s.nblocks == int(fx["nblocks"])
    for ib in range(s.nblocks):
        assert np.array_equal(s.gc_iblocks(ib), fx["gcib"][ib].astype(np.int32))


def _leaves_dfs(s):
    return [s.block_info(ib) for ib in range(s.nblocks)]


def _coordconv(p):
    r, t, ph = p
    st, ct, sp, cp = np.sin(t), np.cos(t), np.sin(ph), np.cos(ph)
    tt = np.arccos(st * sp)
    a = np.arccos(-((st * cp) / np.sqrt((st * st) * (cp * cp) + ct * ct)))
    return np.array([r, tt, np.copysign(abs(a), ct)])


def test_donors_by_brute_force_point_location():
    """Every ghost's donor is the first leaf, in label order, whose box contains
    the ghost centre (Yin panel tried first); -777 only outside the radial range."""
    lay = configs.small_mixed()
    s = Shell(lay)
    s.build_tables()
    info = _leaves_dfs(s)
    ng = 2
    for ib in range(0, s.nblocks, 3):
        rng, fl, (xi, xj, xk) = info[ib]
        gc = s.gc_iblocks(ib)
        NK, NJ, NI = gc.shape
        for K in range(NK):
            for J in range(NJ):
                for I in range(NI):
                    interior = ng <= I < NI - ng and ng <= J < NJ - ng and ng <= K < NK - ng
                    if interior:
                        continue
                    p = np.array([xi[I], xj[J], xk[K]])
                    q = _coordconv(p)
                    found = -777
                    for jb, (r2, f2, _) in enumerate(info):
                        h = p if f2[0] == fl[0] else q
                        lo, hi = r2[:3], r2[3:]
                        if np.all((h - lo) * (h - hi) <= 0.0):
                            found = jb + 1
                            break
                    assert gc[K, J, I] == found, (ib, I, J, K)


def test_entries_weights_and_bases_in_range():
    lay = configs.small_mixed()
    s = Shell(lay)
    s.build_tables()
    e = s.gc_entries()
    n = np.array([lay["ni"], lay["nj"], lay["nk"]])
    assert np.all(e["bijk"] >= 1) and np.all(e["bijk"] <= n)
    assert np.all(e["w"] >= 0.0) and np.all(e["w"] <= 1.0)
    # a clamped base at the top index carries an exactly-zero weight (donor ghosts never contribute)
    top = e["bijk"] == n
    assert np.all(e["w"][top] == 0.0)
    # Yin-Yang rotation is orthogonal: a^2 + b^2 = 1
    yy = e["yy"] == 1
    assert yy.any()
    a, b = e["rot"][yy, 0], e["rot"][yy, 1]
    assert np.max(np.abs(a * a + b * b - 1.0)) < 1e-11
    # counts conserved: every receiving block's entries equal its targets' nGC sum
    for ib in range(s.nblocks):
        assert np.sum(e["recv"] == ib) == sum(t[1].shape[1] for t in s.gc_targets(ib))


def _pull_restated(lay, e):
    """The FP32 pull table restated from the receiver's side: direction s of
    block R (s_a = -1 / +1 for ghosts below / above the interior along axis a,
    0 inside) names donor D when every ghost of R in that region is a near
    copy (same panel, weights within 1e-12 of 0 or 1) from D's interior cell at
    the ghost index minus s * n, and the region is complete."""
    ng = lay.get("ng", 2)
    n = np.array([lay["ni"], lay["nj"], lay["nk"]])
    w = e["w"]
    near = (e["yy"] == 0) & np.all((w < 1e-12) | (w > 1 - 1e-12), axis=1)
    g = e["gijk"] + ng - 1
    p = e["bijk"] + ng - 1 + (w > 0.5)
    side = np.where(g < ng, -1, np.where(g >= ng + n, 1, 0))
    ok = near & np.all(p == g - side * n, axis=1) & np.all((p >= ng) & (p < ng + n), axis=1)
    code = (side[:, 0] + 1) * 9 + (side[:, 1] + 1) * 3 + (side[:, 2] + 1)
    nb = int(e["recv"].max()) + 1
    pull = -np.ones((nb, 27), np.int64)
    fits = bool(np.all(n >= 2 * ng))
    regions = {}
    for q in range(len(g)):
        regions.setdefault((int(e["recv"][q]), int(code[q])), []).append(q)
    n_pulled = 0
    for (R, c), qs in regions.items():
        s = np.array([c // 9 - 1, (c // 3) % 3 - 1, c % 3 - 1])
        size = int(np.prod(np.where(s != 0, ng, n)))
        donors = {int(e["donor"][q]) for q in qs}
        if fits and len(qs) == size and all(ok[q] for q in qs) and len(donors) == 1:
            pull[R, c] = donors.pop()
            n_pulled += len(qs)
    return pull, int(near.sum()), int(near.sum()) - n_pulled


@pytest.mark.parametrize("name", ["mixed", "cubic8"])
def test_ghost_pull_table_restated(name):
    """The library's pull table (built on the host from the donor side) equals
    the receiver-side restatement, entry for entry, with the same copy and
    left-over counts; the 16^3 cubic shell pulls most copies."""
    lay = {"mixed": configs.small_mixed(), "cubic8": configs.cubic((16, 32, 96), 8)}[name]
    s = Shell(lay)
    s.build_tables()
    pull, n_copy, n_rest = s.ghost_pull_table()
    ref, ref_copy, ref_rest = _pull_restated(lay, s.gc_entries())
    assert np.array_equal(pull, ref)
    assert (n_copy, n_rest) == (ref_copy, ref_rest)
    assert (pull >= 0).any() and n_rest < n_copy
