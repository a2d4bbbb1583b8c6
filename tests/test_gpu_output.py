"""AllCells output on the GPU against the reference's own writer.

tests/golden/allcells_mixed.json holds the SHA-256 of every file that the
reference's ModSave_AllCells (3_6/ModSavePlot.f90:520-650, compiled from the
reference sources into the flang harness) wrote for the exchanged initial
state of the small mixed layout (tests/golden/make_allcells_fixture.py).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import make_allcells_fixture as fx
from falcon_amd import Shell

pytestmark = pytest.mark.gpu


def _shell():
    s = Shell(fx.LAYOUT)
    s.set_gc_all()
    nyin = sum(1 for ib in range(s.nblocks) if s.block_info(ib)[1][0])
    bgs, sts = fx.inputs(lambda ib: s.block_info(ib)[2][0], s.pad, s.nblocks, nyin)
    for ib in range(s.nblocks):
        s.set_background(ib, bgs[ib])
        s.set_state(ib, sts[ib])
    # the reference driver's two initial fills (3_6/test1.f90:143-144)
    s.sendrecv_gc(0)
    s.sendrecv_gc(0)
    return s


def test_allcells_files_equal_the_reference_writer(tmp_path):
    print("[allcells] layout=mixed", flush=True)
    if not os.path.exists(fx.OUT):
        pytest.skip("tests/golden/allcells_mixed.json not generated (make_allcells_fixture.py)")
    with open(fx.OUT) as f:
        want = json.load(f)
    s = _shell()
    assert s.nblocks == want["nblocks"]
    d = s.save_allcells(0, str(tmp_path))
    got = sorted(os.listdir(d))
    assert got == sorted(want["sha256"])
    assert open(os.path.join(d, "info.dat")).read() == want["info"]
    bad = [f for f in got if hashlib.sha256(open(os.path.join(d, f), "rb").read()).hexdigest() != want["sha256"][f]]
    assert not bad, bad[:5]
    s.close()


def test_allcells_merged_file(tmp_path):
    """merged = 1: the header, the block files in label order and the brick
    connectivity of merge_AllCells.py in one file; each brick's 8 corners are
    its cell's corners (checked on the corner positions: radius grows from the
    i face to the i+1 face, and the i+1 corners sit one line further)."""
    print("[allcells merged] layout=mixed", flush=True)
    s = _shell()
    d1 = s.save_allcells(3, str(tmp_path / "blocks"))
    d2 = s.save_allcells(3, str(tmp_path / "merged"), merged=True)
    s.close()
    assert sorted(os.listdir(d2)) == [".merged", "info.dat", "merged.dat"]
    text = open(os.path.join(d2, "merged.dat")).read().split("\n")
    info = open(os.path.join(d1, "info.dat")).read().split("\n")
    ni, nj, nk = fx.LAYOUT["ni"], fx.LAYOUT["nj"], fx.LAYOUT["nk"]
    nb = int(info[0].split("=")[1])
    ncorner, ncell = nb * (ni + 1) * (nj + 1) * (nk + 1), nb * ni * nj * nk
    assert text[0] == 'TITLE = "AllCells"'
    assert text[1] == info[5]
    assert text[2] == f"ZONE N={ncorner}, E={ncell}, F=FEPOINT, ET=BRICK"
    blocks = "".join(open(os.path.join(d1, "Block_%08d.dat" % (b + 1))).read() for b in range(nb)).split("\n")[:-1]
    assert text[3:3 + ncorner] == blocks
    xyz = np.array([[float(ln[16 * q:16 * q + 16]) for q in range(3)] for ln in blocks])
    rad = np.linalg.norm(xyz, axis=1)
    conn = np.array([[int(x) for x in ln.split()] for ln in text[3 + ncorner:3 + ncorner + ncell]]) - 1
    assert conn.shape == (ncell, 8) and text[3 + ncorner + ncell] == ""
    assert np.all(rad[conn[:, [1, 2, 5, 6]]] > rad[conn[:, [0, 3, 4, 7]]])
    assert np.all(conn[:, 1] == conn[:, 0] + 1) and np.all(conn[:, 4] - conn[:, 0] == (ni + 1) * (nj + 1))
    assert len(np.unique(conn)) == ncorner
