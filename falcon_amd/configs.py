"""Workload layouts.

Each layout is a dict understood by falcon_amd.Shell and the oracle:
("split", (nr, nt, np)) divides every leaf into nr x nt x np children with the
reference's split rule; ("amr", flags, if_global, rng6_rsun) is one
ModAMR_Divide_OneBranch pass (3_6/ModAMR.f90:65-99); ("refine", ordinals)
splits the listed leaves (0-based, depth-first order) 2x2x2.
"""
import numpy as np

R_RANGE = (0.7, 1.0)
AMR_ALL_R = (0.6, 0.0, 0.0, 1.1, 0.0, 0.0)   # strictly contains both radial ends


def cubic(cells_per_panel, block):
    """Yin-Yang shell with `cells_per_panel` = (nr, nt, np) cells per panel in cubic blocks."""
    nr, nt, npp = cells_per_panel
    assert nr % block == 0 and nt % block == 0 and npp % block == 0
    return {"ni": block, "nj": block, "nk": block, "ng": 2, "r_rsun": R_RANGE,
            "ops": [("split", (nr // block, nt // block, npp // block))]}


# shell of 24x48x144 cells per panel in 8^3 blocks (plain method, CPU-runnable)
SHELL_24x48x144_B8 = cubic((24, 48, 144), 8)
# 64x128x384 per panel in 16^3 blocks (plain FP64 GPU path)
SHELL_64x128x384_B16 = cubic((64, 128, 384), 16)
# 128x256x768 per panel in 16^3 blocks (~50M cells; precision + batching)
SHELL_128x256x768_B16 = cubic((128, 256, 768), 16)

# reference-expressible equivalents (binary splits only): 24x6x18 blocks with
# three global theta/phi levels = 24x48x144 cells per panel, 128 blocks
SHELL_REF_24x48x144 = {"ni": 24, "nj": 6, "nk": 18, "ng": 2, "r_rsun": R_RANGE,
                       "ops": [("amr", (0, 1, 1), 1, AMR_ALL_R)] * 3}


def small_mixed():
    """Small layout with radial and angular splits plus a local refinement on Yin,
    exercising same-level, coarse/fine and Yin-Yang ghosts."""
    return {"ni": 8, "nj": 6, "nk": 12, "ng": 2, "r_rsun": R_RANGE,
            "ops": [("amr", (0, 1, 1), 1, AMR_ALL_R),
                    ("amr", (1, 0, 0), 1, AMR_ALL_R),
                    ("amr", (0, 1, 1), 0, (0.8, 1.2, -0.5, 0.95, 1.6, 0.3))]}


def cubic_refined(cells_per_panel, block, region=(0.8, 1.2, -0.5, 0.95, 1.6, 0.3)):
    """`cubic` plus one local 2x2x2 refinement of the Yin leaves in `region`
    (rng6 in R_sun / radians): cubic blocks at two levels, coarse/fine ghosts."""
    lay = cubic(cells_per_panel, block)
    lay["ops"] = lay["ops"] + [("amr", (1, 1, 1), 0, region)]
    return lay


def cubic_random_amr(cells_per_panel, block, frac1=0.20, frac2=0.05, seed=12345):
    """`cubic` plus seeded local refinement: round(frac1 * N0) of the N0
    level-0 blocks split to level 1, then round(frac2 * N0) of the level-1
    blocks split to level 2 (leaves chosen without replacement)."""
    lay = cubic(cells_per_panel, block)
    nr, nt, npp = cells_per_panel
    n0 = 2 * (nr // block) * (nt // block) * (npp // block)
    rng = np.random.default_rng(seed)
    levels = [0] * n0
    first = sorted(rng.choice(n0, size=int(round(frac1 * n0)), replace=False).tolist())
    new = []
    for i, lv in enumerate(levels):
        new += [1] * 8 if i in set(first) else [lv]
    levels = new
    lvl1 = [i for i, lv in enumerate(levels) if lv == 1]
    second = sorted(rng.choice(lvl1, size=int(round(frac2 * n0)), replace=False).tolist())
    lay["ops"] = lay["ops"] + [("refine", first), ("refine", second)]
    return lay


# learned-FCG workload: level-0 32x64x192 per panel in 16^3 blocks, 20% / 5% to levels 1 / 2
SHELL_AMR_32x64x192_B16 = cubic_random_amr((32, 64, 192), 16)
