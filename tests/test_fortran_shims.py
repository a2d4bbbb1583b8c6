"""The Fortran drop-in layer (falcon_amd/fortran) compiles against the
reference's own module files: the shim modules keep the reference's module and
entry-point names (ModAdvance_rk4, ModCommunication_SetGCAll,
ModCommunication_SendRecvGC_new) and types (YYTree, BlockType)."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
REFMODS = os.path.join(REPO, "oracle", "_ref")


@pytest.mark.skipif(not (os.path.exists(FLANG) and os.path.exists(os.path.join(REFMODS, ".mods_done"))),
                    reason="flang or the reference's module files (make -C oracle ref) are missing")
def test_shims_compile_against_reference_modules(tmp_path):
    src = os.path.join(REPO, "falcon_amd", "fortran")
    work = tmp_path / "fortran"
    shutil.copytree(src, work, ignore=shutil.ignore_patterns("_build"))
    r = subprocess.run(["make", "-s", "-C", str(work), f"REFMODS={REFMODS}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    for m in ("modadvance.mod", "modcommunication.mod", "modheatconduction.mod", "modpfss.mod", "falcon_c.mod"):
        assert (work / "_build" / m).exists(), m
