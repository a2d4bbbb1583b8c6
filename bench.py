"""Benchmark on one MI355X: RK4 cell-updates/s (FP32-deviation vs FP64) and
heat-solve seconds to 1e-10.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--legs rk,heat,learned] [--no-cpu-baseline] [--no-autotune]

A "step" is one full RK4 step (4 RHS stages, artificial diffusion, 5 ghost
fills) over every block of the workload; all blocks are resident in HBM when
the timed region starts.  `value` is the FP32-deviation path's cell-updates/s
on the 128x256x768-per-panel shell; the FP64 path, the stage kernel's
roofline, the heat leg and the CPU baselines ride in the same JSON line.
Before each precision is timed, the autotuner picks the (theta, phi) tile edge
of the stage and diffusion marches from timings on the same state ("tiles").
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from falcon_amd import HeatSolver, Pfss, Shell, configs  # noqa: E402

WORKLOADS = {
    "shell_128x256x768_b16": configs.SHELL_128x256x768_B16,
    "shell_64x128x384_b16": configs.SHELL_64x128x384_B16,
    "shell_24x48x144_b8": configs.SHELL_24x48x144_B8,
}

# algorithmic bytes per interior cell per stage-kernel launch (compulsory
# interior traffic: read the stage input, read the accumulation base, write the
# output).  FP64: 64 + 64 + 64.  FP32 path: stages 1-3 read the FP32 stage input
# and the FP32 copy of U^n and write FP32 (32 + 32 + 32); stage 4 reads the FP32
# input and FP64 U^n and writes FP64 U^n+1 and its FP32 copy (32 + 64 + 64 + 32)
STAGE_BYTES = {"fp64": [192, 192, 192, 192], "fp32": [96, 96, 96, 192]}
HBM_PEAK_GBS = 8000.0


def run_precision(shell, prec, steps, warmup, cfl, autotune=True):
    shell.set_precision(prec)
    shell.init_synthetic(seed=12345, s_amp=1e-4, b_ratio=1e-3)
    # tile edges of the stage and diffusion marches, timed on this state
    # (outside the timed region; results do not depend on the edges)
    tune = shell.autotune(reps=5) if autotune else None
    shell.advance(warmup, cfl, graph=True)
    shell.synchronize()
    t0 = time.perf_counter()
    shell.advance(steps, cfl, graph=True)
    shell.synchronize()
    el = time.perf_counter() - t0
    prof = shell.profile_step(3, cfl)
    cells = shell.interior_cells
    stage_bytes = float(np.mean(STAGE_BYTES[prec])) * cells
    achieved = stage_bytes / (prof["stage"] * 1e-3) / 1e9
    return {
        "cell_updates_per_s": cells * steps / el,
        "ms_per_step": 1e3 * el / steps,
        "tiles": tune,
        "kernel_ms": prof,
        "stage_roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="shell_128x256x768_b16", choices=sorted(WORKLOADS))
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the 16-edge march tiles instead of timing both")
    ap.add_argument("--learned-sweeps", default="2,2",
                    help="finest-level pre,post smoothing sweeps of the V-cycle under the learned correction")
    ap.add_argument("--learned-pairs", type=int, default=4000, help="solver-generated residual/error snapshots")
    ap.add_argument("--learned-blocks", type=int, default=16, help="block images recorded per snapshot")
    ap.add_argument("--learned-epochs", type=int, default=60)
    ap.add_argument("--learned-blockstep", type=int, default=1, help="1: correction step per 16^3 block, 0: per system")
    ap.add_argument("--learned-loss-h1", type=float, default=0.01,
                    help="training loss: < 0 mean squared error; >= 0 gradient-weighted error with this mass weight")
    ap.add_argument("--learned-loss-w", default="1,1,1",
                    help="face weights (r, theta, phi) of the gradient-weighted training loss")
    ap.add_argument("--learned-loss-scale-free", type=int, default=0,
                    help="1: train on the scale-free form of the gradient-weighted loss")
    ap.add_argument("--learned-iters", type=int, default=0,
                    help="directions built with the learned correction before the plain V-cycle takes over (0: all)")
    ap.add_argument("--legs", default="rk,heat,learned,pfss,glm",
                    help="comma list of legs: rk (RK4 FP64/FP32), heat (MG-PCG), learned (learned FCG vs MG-PCG), "
                         "pfss (potential-field solve), glm (FP64 step with GLM divergence cleaning)")
    ap.add_argument("--error-steps", type=int, default=100,
                    help="also step the FP64 and FP32 paths this many steps from the same state and report the "
                         "FP32 path's per-variable difference (0: skip)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run a device-to-device copy of this many GiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    legs = set(args.legs.split(","))
    lay = WORKLOADS[args.workload]
    res, setup_s, nblocks, ncells = None, None, None, None
    if "rk" in legs:
        t0 = time.perf_counter()
        shell = Shell(lay)
        shell.background_polytrope(20.0)
        shell.set_gc_all()
        setup_s = time.perf_counter() - t0
        nblocks, ncells = shell.nblocks, shell.interior_cells
        print(f"# setup {args.workload}: {nblocks} blocks, {ncells} cells, {setup_s:.1f} s", file=sys.stderr, flush=True)
        res = {}
        for prec in ("fp64", "fp32"):
            res[prec] = run_precision(shell, prec, args.steps, args.warmup, args.cfl, not args.no_autotune)
            print(f"# {prec}: {res[prec]['cell_updates_per_s']:.4e} cell-updates/s, "
                  f"{res[prec]['ms_per_step']:.3f} ms/step, kernels {res[prec]['kernel_ms']}", file=sys.stderr,
                  flush=True)
        shell.close()
        if args.error_steps > 0:
            err = fp32_error(lay, args.error_steps, args.cfl)
            res["fp32_error"] = err
            print(f"# fp32 vs fp64 after {args.error_steps} steps: {json.dumps(err)}", file=sys.stderr, flush=True)
        traffic = load_traffic(args.workload)
        for prec in ("fp32", "fp64"):
            res[prec]["stage_roofline"]["traffic"] = traffic.get(prec, {}).get("stage_bytes_per_launch")

    heat = None
    if "heat" in legs:
        heat = run_heat(cpu_baseline=not args.no_cpu_baseline)
        print(f"# heat: {heat['mg_pcg']}", file=sys.stderr, flush=True)
    if "learned" in legs:
        learned = run_learned(npairs=args.learned_pairs, blocks_per_pair=args.learned_blocks,
                              epochs=args.learned_epochs, sweeps=tuple(int(v) for v in args.learned_sweeps.split(",")),
                              blockstep=bool(args.learned_blockstep), learned_iters=args.learned_iters,
                              loss_h1=args.learned_loss_h1,
                              loss_w=tuple(float(v) for v in args.learned_loss_w.split(",")),
                              loss_scale_free=bool(args.learned_loss_scale_free))
        print(f"# learned: {json.dumps(learned)}", file=sys.stderr, flush=True)
        heat = heat or {}
        heat["learned_fcg"] = learned

    extras = {}
    if "pfss" in legs:
        extras["pfss"] = run_pfss()
        print(f"# pfss: {json.dumps(extras['pfss'])}", file=sys.stderr, flush=True)
    if "glm" in legs:
        extras["glm"] = run_glm(args.workload, args.steps, args.warmup, args.cfl)
        if res:
            extras["glm"]["fp64_ms_per_step_without"] = res["fp64"]["ms_per_step"]
        print(f"# glm: {json.dumps(extras['glm'])}", file=sys.stderr, flush=True)

    cpu = None
    if "rk" in legs and not args.no_cpu_baseline:
        cpu = run_cpu_baseline()

    head = res["fp32"] if res else None
    out = {
        "metric": "RK4 cell-updates/s FP32-dev vs FP64; heat-solve s @1e-10 learned FCG vs MG-PCG",
        "value": head["cell_updates_per_s"] if head else None,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"] if head else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencils / f64 state",
        "data": "synthetic (seeded polytrope shell, entropy noise, curl-A field; seeded Gaussian T')",
        "config": {"workload": args.workload, "blocks": nblocks, "block": [lay["ni"], lay["nj"], lay["nk"]],
                   "interior_cells": ncells, "cfl": args.cfl},
        "fp64": ({"value": res["fp64"]["cell_updates_per_s"], "ms_per_step": res["fp64"]["ms_per_step"],
                  "kernel_ms": res["fp64"]["kernel_ms"], "stage_roofline": res["fp64"]["stage_roofline"],
                  "tiles": res["fp64"]["tiles"]} if res else None),
        "fp32_over_fp64": head["cell_updates_per_s"] / res["fp64"]["cell_updates_per_s"] if head else None,
        "fp32_error": res.get("fp32_error") if res else None,
        "kernel_ms": head["kernel_ms"] if head else None,
        "tiles": head["tiles"] if head else None,
        "roofline": head["stage_roofline"] if head else None,
        "setup_s": setup_s,
        "cpu_baseline": cpu,
        "heat": heat,
        "extras": extras or None,
    }
    print(json.dumps(out))


def fp32_error(lay, steps, cfl):
    """FP32-deviation path against the FP64 path after `steps` steps from the
    same seeded state, two ways: each path with its own CFL dt (the metric as
    stated), and the FP32 path stepped with the FP64 path's dt sequence (the
    arithmetic error alone).  Per variable: max|a-b| / max|b| over the interior
    cells, the cells above 1e-6 of max|b|, and where the largest difference is
    (block, i, j, k).  The differences are formed on the device."""
    def make(prec):
        s = Shell(lay)
        s.background_polytrope(20.0)
        s.set_gc_all()
        s.set_precision(prec)
        s.init_synthetic(seed=12345, s_amp=1e-4, b_ratio=1e-3)
        return s

    def summary(a, ref):
        d = a.state_diff(ref, 1e-6)
        rel = d["max_diff"] / np.where(d["max_ref"] > 0, d["max_ref"], 1.0)
        return {"max_rel": [float("%.4g" % x) for x in rel], "cells_above_1e-6": [int(x) for x in d["n_above"]],
                "argmax": d["argmax"]}

    ref = make("fp64")
    dts = [ref.advance(1, cfl, graph=False) for _ in range(steps)]
    ref.synchronize()
    own = make("fp32")
    own.advance(steps, cfl, graph=True)
    own.synchronize()
    res = {"steps": steps, "cells": ref.interior_cells, "own_dt": summary(own, ref)}
    own.close()
    shared = make("fp32")
    for dt in dts:
        shared.set_fixed_dt(dt)
        shared.advance(1, cfl, graph=False)
    shared.synchronize()
    res["shared_dt"] = summary(shared, ref)
    shared.close()
    ref.close()
    res["max_rel"] = [max(a, b) for a, b in zip(res["own_dt"]["max_rel"], res["shared_dt"]["max_rel"])]
    return res


def copy_calibration(gib):
    """A plain copy of known size through torch: its FETCH_SIZE/WRITE_SIZE per
    dispatch give the scale between the counters and bytes moved."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4}))


def run_cpu_baseline():
    """Bounded CPU sample of the same RK4 method on the host cores.

    Preferred: the reference's own Fortran (ModAdvance_rk4, ModCommunication
    and the numerics modules, compiled from the reference sources by
    oracle/Makefile, at -O3 as the reference's own Makefile builds it) stepping
    the 64x128x384-per-panel shell in the reference's block layout, under
    MPICH's mpirun on up to 16 host cores, with the one-process rate beside it.
    Fallback when no harness binary is present: the FP64 C oracle with OpenMP."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    import pyoracle
    import subprocess
    fast = os.path.join(REPO, "oracle", "_ref", "fast", "falcon_ref_harness")
    for exe, flags in ((fast, "-O3"), (pyoracle.HARNESS, "-O2 -ffp-contract=off")):
        if os.path.exists(exe):
            try:
                return _cpu_baseline_reference(exe, flags)
            except (subprocess.CalledProcessError, OSError) as e:
                print(f"# reference harness {exe} failed ({e}); next baseline", file=sys.stderr, flush=True)
    return _cpu_baseline_port(pyoracle)


# the 64x128x384-per-panel shell as the reference's tree expresses it: blocks
# of 64x16x48 cells, three global theta/phi levels, 128 blocks
CPU_LAYOUT = {"ni": 64, "nj": 16, "nk": 48, "ng": 2, "r_rsun": (0.7, 1.0),
              "ops": [("amr", (0, 1, 1), 1, configs.AMR_ALL_R)] * 3}


def _cpu_baseline_reference(exe, flags):
    import shutil
    import subprocess
    import tempfile
    import ref_io
    lay = CPU_LAYOUT
    s = Shell(lay)
    s.background_polytrope(20.0)
    s.set_gc_all()
    s.init_synthetic(seed=12345, s_amp=1e-4, b_ratio=1e-3)
    bgs = [s.get_background(ib) for ib in range(s.nblocks)]
    sts = [s.get_state(ib) for ib in range(s.nblocks)]
    nb = s.nblocks
    s.close()
    cells = nb * lay["ni"] * lay["nj"] * lay["nk"]
    # the host cores of one GPU's share of the box
    ranks = max(1, min(16, os.cpu_count() or 1))
    mpirun = os.path.join(os.environ.get("MPIDIR", "/opt/conda"), "bin", "mpirun")
    if not os.path.exists(mpirun):
        mpirun = shutil.which("mpirun")

    ref_io.raise_stack_limit()

    def timed(nsteps, launcher):
        with tempfile.TemporaryDirectory() as td:
            fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            ref_io.write_input(fi, lay, 0.5, nsteps, [], 3, bgs, sts)
            out = subprocess.run(list(launcher) + [exe, fi, fo], check=True, capture_output=True, text=True).stdout
        line = [ln for ln in out.splitlines() if ln.startswith("harness_steps")][-1].split()
        return float(line[3])

    n1 = 2
    el1 = timed(n1, [])
    one = {"value": cells * n1 / el1, "cores": 1, "sample": f"{n1} steps, {el1:.1f} s"}
    res = {"unit": "cell-updates/s", "kind": "reference", "one_core": one}
    nn = 20
    try:
        if ranks < 2 or not mpirun:
            raise RuntimeError("no mpirun or a single host core")
        el = timed(nn, [mpirun, "-n", str(ranks)])
        res.update({"value": cells * nn / el, "cores": ranks,
                    "sample": f"reference Fortran RK4 (ModAdvance_rk4 + ModCommunication, flang {flags}, MPICH "
                              f"mpirun -n {ranks}), 64x128x384-per-panel shell in {nb} blocks of 64x16x48, "
                              f"{nn} steps, {el:.1f} s; one process: {one['value']:.4g} cell-updates/s"})
    except (RuntimeError, subprocess.CalledProcessError, OSError) as e:
        res.update({"value": one["value"], "cores": 1,
                    "sample": f"reference Fortran RK4 (flang {flags}, one process; the multi-rank run failed: "
                              f"{type(e).__name__}), 64x128x384-per-panel shell in {nb} blocks of 64x16x48, "
                              f"{n1} steps, {el1:.1f} s"})
    return res


def _cpu_baseline_port(pyoracle):
    lay = configs.cubic((32, 64, 192), 16)
    threads = min(os.cpu_count() or 1, 16)
    pyoracle.set_threads(threads)
    o = pyoracle.Oracle(lay)
    s = Shell(lay)
    s.background_polytrope(20.0)
    s.set_gc_all()
    s.init_synthetic(seed=12345)
    for ib in range(o.nblocks):
        o.set_background(ib, s.get_background(ib))
        o.set_state(ib, s.get_state(ib))
    o.step(0.5)  # warm caches
    n, t0 = 0, time.perf_counter()
    while True:
        o.step(0.5)
        n += 1
        if time.perf_counter() - t0 > 10.0 or n >= 20:
            break
    el = time.perf_counter() - t0
    cells = o.nblocks * 16 ** 3
    return {"value": cells * n / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 RK4, 32x64x192-per-panel shell in 16^3 blocks ({o.nblocks} blocks), {n} steps, {el:.1f} s"}


HEAT_WORKLOAD = "heat_shell_64x128x384_b16"


def run_heat(cpu_baseline=True, repeats=3):
    """Backward-Euler heat solve to 1e-10: 64x128x384 cells per panel in 16^3
    blocks, kappa = kappa0 (r/0.7 R_sun)^-3, dt = 100 x explicit limit, seeded
    Gaussian T' (sigma 1e-3); MG-preconditioned FCG in FP64."""
    lay = configs.SHELL_64x128x384_B16
    s = Shell(lay)
    hs = HeatSolver(s, 1.0, 0.7, 1)
    dt = 100.0 * hs.dt_explicit
    hs.set_dt(dt)
    T = 1e-3 * np.random.default_rng(12345).standard_normal(hs.shape)
    hs.set_T(T)
    hs.solve(1e-10, 300)                    # warm-up (first launches)
    runs = [hs.solve(1e-10, 300) for _ in range(repeats)]
    secs = sorted(r[2] for r in runs)
    it, rel = int(runs[0][0][0]), float(runs[0][1][0])
    out = {"workload": HEAT_WORKLOAD, "cells": hs.cells, "levels": hs.levels, "dt_over_explicit": 100.0,
           "tol": 1e-10,
           "mg_pcg": {"seconds": secs[len(secs) // 2], "iterations": it, "relres": rel},
           "learned_fcg": None}
    if cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(lay)
        H = pyoracle.HeatOracle(o, 1.0, 0.7)
        t0 = time.perf_counter()
        _, ito, relo = H.solve(dt, T[0], 1e-10, 300)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": el, "unit": "s", "cores": 1, "kind": "port", "iterations": ito,
                               "sample": "oracle FP64 MG-preconditioned FCG, same system, one solve to 1e-10"}
    return out


def run_pfss(repeats=3):
    """PFSS potential field (3_5/ModPFSS.f90) on the 64x128x384-per-panel shell
    (1536 blocks of 16^3): magnetogram of a centred dipole plus an offset one,
    MG-preconditioned FCG to 1e-10, then B = -grad(phi)."""
    s = Shell(configs.SHELL_64x128x384_B16)
    pf = Pfss(s)
    r0 = 0.7 * float(np.float32(6.955e10))
    xyz = np.array([[0.0, 0.0, 0.0], [0.3 * r0, -0.2 * r0, 0.5 * r0]])
    q = np.array([1.0, -0.4])
    l = np.array([[0.0, 0.0, 0.05 * r0], [0.02 * r0, 0.03 * r0, 0.0]])
    pf.set_magnetogram(pf.dipole_magnetogram(xyz, q, l))
    pf.solve(1e-10, 500)                    # warm-up (first launches)
    runs = [pf.solve(1e-10, 500) for _ in range(repeats)]
    secs = sorted(r[2] for r in runs)
    t0 = time.perf_counter()
    pf.field()
    field_s = time.perf_counter() - t0
    s.close()
    return {"workload": "pfss_shell_64x128x384_b16", "cells": int(np.prod(pf.shape)), "tol": 1e-10,
            "seconds": secs[len(secs) // 2], "iterations": runs[0][0], "relres": runs[0][1],
            "field_seconds_with_transfer": field_s}


def run_glm(workload, steps, warmup, cfl):
    """FP64 step with GLM divergence cleaning (DivB_option=1) on the RK
    workload: the same seeded state, psi = 0 at the start."""
    s = Shell(WORKLOADS[workload])
    s.background_polytrope(20.0)
    s.set_gc_all()
    s.set_precision("fp64")
    s.set_divb(1)
    s.init_synthetic(seed=12345, s_amp=1e-4, b_ratio=1e-3)
    s.advance(warmup, cfl, graph=True)
    s.synchronize()
    t0 = time.perf_counter()
    s.advance(steps, cfl, graph=True)
    s.synchronize()
    el = time.perf_counter() - t0
    cells = s.interior_cells
    s.close()
    return {"workload": workload, "cell_updates_per_s": cells * steps / el, "ms_per_step": 1e3 * el / steps}


LEARNED_WORKLOAD = "heat_amr_32x64x192_b16_x32"


def run_learned(npairs=4000, blocks_per_pair=16, epochs=60, batch=128, lr=2e-3, repeats=3, sweeps=(2, 2), blockstep=True,
                learned_iters=0, loss_h1=0.01, loss_w=(1.0, 1.0, 1.0), loss_scale_free=False):
    """32 concurrent backward-Euler solves on the locally refined shell
    (level-0 32x64x192 per panel in 16^3 blocks, seeded 20% of blocks to level
    1 and 5% to level 2), dt spread geometrically over 10-300 x the explicit
    limit, seeded Gaussian T' (sigma 1e-3), to 1e-10: plain MG-FCG against
    FCG with the learned correction, trained first on `npairs` block-image
    pairs generated by the solver's own systems (training time reported
    separately, not in the solve time).  The learned correction builds the
    first `learned_iters` search directions, the plain V-cycle the rest
    (0: all).  loss_h1 >= 0 trains on the gradient-weighted error with that
    mass weight, < 0 on the mean squared error."""
    lay = configs.SHELL_AMR_32x64x192_B16
    nsys = 32
    s = Shell(lay)
    hs = HeatSolver(s, 1.0, 0.7, nsys)
    hs.set_learned_sweeps(*sweeps)
    hs.set_learned_blockstep(blockstep)
    hs.set_learned_iters(learned_iters)
    hs.set_training_loss(loss_h1, loss_w, loss_scale_free)
    loss, pairs, train_s = hs.train_corrector(npairs=npairs, blocks_per_pair=blocks_per_pair, dt_range=(10.0, 300.0),
                                              max_solves=20, epochs=epochs, batch=batch, lr=lr, seed=12345)
    dts = hs.dt_explicit * 10.0 * 30.0 ** (np.arange(nsys) / (nsys - 1))
    T = 1e-3 * np.random.default_rng(54321).standard_normal(hs.shape)
    hs.set_dt(dts)
    hs.set_T(T)
    out = {"workload": LEARNED_WORKLOAD, "blocks": s.nblocks, "systems": nsys, "cells_per_system": hs.cells,
           "tol": 1e-10, "dt_over_explicit": [10.0, 300.0],
           "learned_finest_sweeps": list(sweeps), "learned_blockstep": bool(blockstep),
           "learned_iters": int(learned_iters),
           "training_loss": "mse" if loss_h1 < 0 else {"gradient_weighted_mass": float(loss_h1),
                                                        "face_weights_r_theta_phi": list(loss_w),
                                                        "scale_free": bool(loss_scale_free)},
           "training": {"pairs": pairs, "blocks_per_pair": blocks_per_pair, "epochs": epochs, "batch": batch, "lr": lr,
                        "final_mse": loss, "seconds": train_s}}
    for name, learned in (("mg_pcg", False), ("learned_fcg", True)):
        hs.solve(1e-10, 300, use_corrector=learned)        # warm-up
        runs = [hs.solve(1e-10, 300, use_corrector=learned) for _ in range(repeats)]
        secs = sorted(r[2] for r in runs)
        it, rel = runs[0][0], runs[0][1]
        out[name] = {"seconds": secs[len(secs) // 2], "iterations_max": int(it.max()),
                     "iterations_mean": float(it.mean()), "relres_max": float(rel.max())}
    out["speedup"] = out["mg_pcg"]["seconds"] / out["learned_fcg"]["seconds"]
    return out


def load_traffic(workload):
    """HBM bytes per launch of the stage kernels from rocprofv3 --pmc runs
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size), stored
    as data in profiles/pmc_traffic.json; None when not measured."""
    path = os.path.join(REPO, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        data = json.load(f)
    return data.get(workload, {})


if __name__ == "__main__":
    main()
