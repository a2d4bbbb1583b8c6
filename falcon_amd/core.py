"""ctypes binding of libfalcon_hip.so and the host-side driver objects.

The library is required: importing a compute entry point without the built
library raises, there is no CPU fallback.  Arrays handed to and from the library
use the reference's Fortran layout; in numpy they are C-contiguous arrays of
shape (8, NK, NJ, NI) for a padded block state (the C view of Fortran
(NI, NJ, NK, 8)) and (8, nk, nj, ni) for an interior field.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libfalcon_hip.so")
R_SUN = float(np.float32(6.955e10))   # 3_6/ModConst.f90:5 (default real)

_lib = None


class FalconError(RuntimeError):
    pass


def load():
    """Load libfalcon_hip.so (built by `python -m falcon_amd.build`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FalconError(f"{LIB_PATH} is missing: build it with `python -m falcon_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, i64p = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_int64)
    fp = C.POINTER(C.c_float)
    sig = {
        "falcon_create": [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double],
        "falcon_destroy": [vp],
        "falcon_tree_divide_all": [vp, C.c_int, C.c_int, C.c_int],
        "falcon_tree_amr": [vp, ip, C.c_int, dp],
        "falcon_tree_set_nodes": [vp, C.c_int, ip, ip, dp],
        "falcon_tree_refine": [vp, C.c_int, ip],
        "falcon_nblocks": [vp, ip, ip],
        "falcon_block_info": [vp, C.c_int, dp, ip, dp, dp, dp],
        "falcon_set_gc_all": [vp],
        "falcon_build_tables": [vp],
        "falcon_gc_iblocks": [vp, C.c_int, ip],
        "falcon_gc_ntargets": [vp, C.c_int, ip],
        "falcon_gc_target": [vp, C.c_int, C.c_int, ip, ip, ip],
        "falcon_gc_entries": [vp, i64p, ip, ip, ip, ip, dp, ip, dp],
        "falcon_ghost_pull_table": [vp, ip, i64p, i64p],
        "falcon_set_background": [vp, C.c_int, dp],
        "falcon_background_polytrope": [vp, C.c_double, C.c_double],
        "falcon_get_background": [vp, C.c_int, dp],
        "falcon_set_state": [vp, C.c_int, dp],
        "falcon_get_state": [vp, C.c_int, dp],
        "falcon_set_divb": [vp, C.c_int],
        "falcon_set_device_tables": [vp, C.c_int],
        "falcon_set_psi": [vp, C.c_int, dp],
        "falcon_get_psi": [vp, C.c_int, dp],
        "falcon_init_synthetic": [vp, C.c_uint64, C.c_double, C.c_double],
        "falcon_set_precision": [vp, C.c_int],
        "falcon_set_ghost_pull": [vp, C.c_int],
        "falcon_set_tiles": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
        "falcon_get_tiles": [vp, C.c_int, ip],
        "falcon_autotune": [vp, C.c_int, ip, dp],
        "falcon_sendrecv_gc": [vp, C.c_int],
        "falcon_advance_rk4": [vp, C.c_double, dp],
        "falcon_advance": [vp, C.c_int, C.c_double, C.c_int, dp],
        "falcon_compute_dt": [vp, C.c_double, dp],
        "falcon_set_fixed_dt": [vp, C.c_double],
        "falcon_check_state": [vp, C.c_int, i64p, ip],
        "falcon_save_allcells": [vp, C.c_int, C.c_char_p, C.c_int],
        "falcon_format_es16": [dp, C.c_int, C.c_char_p],
        "falcon_compute_dt_noalloc": [vp, C.c_double, dp],
        "falcon_state_diff": [vp, vp, C.c_double, dp, dp, i64p, i64p],
        "falcon_rhs": [vp, dp],
        "falcon_diffusion_rhs": [vp, dp],
        "falcon_profile_step": [vp, C.c_int, C.c_double, dp],
        "falcon_device_bytes": [vp, i64p],
        "falcon_synchronize": [vp],
        "falcon_heat_setup": [vp, C.c_double, C.c_double, C.c_int],
        "falcon_heat_info": [vp, ip, i64p, dp],
        "falcon_heat_set_dt": [vp, dp],
        "falcon_heat_set_T": [vp, dp],
        "falcon_heat_seed_T": [vp, C.c_uint64, C.c_double],
        "falcon_heat_solve": [vp, C.c_double, C.c_int, C.c_int, ip, dp, dp],
        "falcon_heat_get_T": [vp, C.c_int, dp],
        "falcon_heat_apply": [vp, dp, dp],
        "falcon_heat_vcycle": [vp, dp, dp],
        "falcon_pfss_setup": [vp],
        "falcon_pfss_bottom_blocks": [vp, ip, ip],
        "falcon_pfss_bottom_areas": [vp, dp],
        "falcon_pfss_set_magnetogram": [vp, dp],
        "falcon_pfss_dipole_magnetogram": [vp, C.c_int, dp, dp, dp, dp],
        "falcon_pfss_solve": [vp, C.c_double, C.c_int, ip, dp, dp],
        "falcon_pfss_get_phi": [vp, dp],
        "falcon_pfss_field": [vp, dp],
        "falcon_cnn_init": [vp, C.c_uint64],
        "falcon_cnn_set_params": [vp, fp],
        "falcon_cnn_get_params": [vp, fp],
        "falcon_cnn_forward": [vp, C.c_int64, fp, fp],
        "falcon_cnn_gradient": [vp, C.c_int, fp, fp, dp],
        "falcon_cnn_train": [vp, C.c_int64, fp, fp, C.c_int, C.c_int, C.c_double, C.c_uint64, dp],
        "falcon_heat_train_corrector": [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                        C.c_double, C.c_uint64, dp, i64p, dp],
        "falcon_heat_attach_corrector": [vp],
        "falcon_heat_set_learned_sweeps": [vp, C.c_int, C.c_int],
        "falcon_heat_set_coarse": [vp, C.c_double, C.c_int],
        "falcon_heat_set_graph": [vp, C.c_int],
        "falcon_heat_set_learned_blockstep": [vp, C.c_int],
        "falcon_heat_set_learned_iters": [vp, C.c_int],
        "falcon_cnn_set_loss": [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int],
        "falcon_last_error": [C.c_char_p, C.c_int],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        buf = C.create_string_buffer(1024)
        _lib.falcon_last_error(buf, 1024)
        raise FalconError(buf.value.decode())


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Shell:
    """One Yin-Yang shell of blocks resident in HBM (one library handle).

    `layout` = {"ni","nj","nk","ng","r_rsun":(lo,hi),"ops":[...]} where each op is
    ("split", (nr,nt,np)) or ("amr", flags3, if_global, rng6_rsun).
    """

    def __init__(self, layout):
        L = load()
        self.L = L
        self.layout = layout
        ni, nj, nk, ng = layout["ni"], layout["nj"], layout["nk"], layout.get("ng", 2)
        self.ni, self.nj, self.nk, self.ng = ni, nj, nk, ng
        self.pad = (ni + 2 * ng, nj + 2 * ng, nk + 2 * ng)
        h = C.c_void_p()
        r0, r1 = layout["r_rsun"]
        _check(L.falcon_create(C.byref(h), ni, nj, nk, ng, r0 * R_SUN, r1 * R_SUN))
        self.h = h
        for op in layout.get("ops", []):
            if op[0] == "split":
                _check(L.falcon_tree_divide_all(h, *[int(x) for x in op[1]]))
            elif op[0] == "refine":
                ids = np.asarray(op[1], dtype=np.int32)
                _check(L.falcon_tree_refine(h, len(ids), _p(ids, C.c_int32)))
            elif op[0] == "amr":
                f = np.asarray(op[1], dtype=np.int32)
                r = np.asarray(op[3], dtype=np.float64)
                _check(L.falcon_tree_amr(h, _p(f, C.c_int32), int(op[2]), _p(r)))
            else:
                raise ValueError(op)
        nb, ny = C.c_int32(), C.c_int32()
        _check(L.falcon_nblocks(h, C.byref(nb), C.byref(ny)))
        self.nblocks, self.nleaf_yin = nb.value, ny.value
        self.on_device = False

    def close(self):
        if getattr(self, "h", None):
            self.L.falcon_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- layout and tables
    def block_info(self, ib):
        rng = np.zeros(6)
        fl = np.zeros(3, dtype=np.int32)
        xi, xj, xk = (np.zeros(n) for n in self.pad)
        _check(self.L.falcon_block_info(self.h, ib, _p(rng), _p(fl, C.c_int32), _p(xi), _p(xj), _p(xk)))
        return rng, fl, (xi, xj, xk)

    def build_tables(self):
        _check(self.L.falcon_build_tables(self.h))

    def set_gc_all(self):
        _check(self.L.falcon_set_gc_all(self.h))
        self.on_device = True

    def set_device_tables(self, on=True):
        """build the ghost tables of set_gc_all on the GPU (donor search,
        targets, interpolation data) instead of the host"""
        _check(self.L.falcon_set_device_tables(self.h, 1 if on else 0))

    def gc_iblocks(self, ib):
        out = np.zeros(self.pad[::-1], dtype=np.int32)
        _check(self.L.falcon_gc_iblocks(self.h, ib, _p(out, C.c_int32)))
        return out

    def gc_targets(self, ib):
        n = C.c_int32()
        _check(self.L.falcon_gc_ntargets(self.h, ib, C.byref(n)))
        res = []
        for it in range(n.value):
            d, g = C.c_int32(), C.c_int32()
            _check(self.L.falcon_gc_target(self.h, ib, it, C.byref(d), C.byref(g), None))
            ijk = np.zeros((3, g.value), dtype=np.int32)
            _check(self.L.falcon_gc_target(self.h, ib, it, C.byref(d), C.byref(g), _p(ijk, C.c_int32)))
            res.append((d.value, ijk))
        return res

    def gc_entries(self):
        n = C.c_int64()
        _check(self.L.falcon_gc_entries(self.h, C.byref(n), None, None, None, None, None, None, None))
        n = n.value
        recv = np.zeros(n, np.int32); donor = np.zeros(n, np.int32)
        gijk = np.zeros((n, 3), np.int32); bijk = np.zeros((n, 3), np.int32)
        w = np.zeros((n, 3)); yy = np.zeros(n, np.int32); rot = np.zeros((n, 2))
        nn = C.c_int64()
        _check(self.L.falcon_gc_entries(self.h, C.byref(nn), _p(recv, C.c_int32), _p(donor, C.c_int32),
                                        _p(gijk, C.c_int32), _p(bijk, C.c_int32), _p(w), _p(yy, C.c_int32), _p(rot)))
        return dict(recv=recv, donor=donor, gijk=gijk, bijk=bijk, w=w, yy=yy, rot=rot)

    def ghost_pull_table(self):
        """FP32 step's pull table (nblocks, 27): the same-level block whose
        interior holds each direction's ghosts, or -1; with the number of FP32
        copy entries and of those the ghost fill still runs."""
        pull = np.zeros((self.nblocks, 27), np.int32)
        nc, nr = C.c_int64(), C.c_int64()
        _check(self.L.falcon_ghost_pull_table(self.h, _p(pull, C.c_int32), C.byref(nc), C.byref(nr)))
        return pull, nc.value, nr.value

    # ---- data
    def set_background(self, ib, bg):
        bg = np.ascontiguousarray(bg, dtype=np.float64)
        _check(self.L.falcon_set_background(self.h, ib, _p(bg)))

    def get_background(self, ib):
        out = np.zeros((8, self.pad[0]))
        _check(self.L.falcon_get_background(self.h, ib, _p(out)))
        return out

    def background_polytrope(self, contrast=20.0, rho_top=1.0e-2):
        _check(self.L.falcon_background_polytrope(self.h, contrast, rho_top))

    def set_state(self, ib, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        _check(self.L.falcon_set_state(self.h, ib, _p(u)))

    def get_state(self, ib):
        out = np.zeros((8,) + self.pad[::-1])
        _check(self.L.falcon_get_state(self.h, ib, _p(out)))
        return out

    def set_divb(self, option):
        """DivB_option (3_6/ModParameters.f90:70): 0 none, 1 GLM divergence
        cleaning with psi as a ninth stepped variable (FP64 path only)."""
        _check(self.L.falcon_set_divb(self.h, int(option)))

    def set_psi(self, ib, psi):
        """padded psi of block ib, (NK, NJ, NI) (the C view of Fortran order)"""
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        if psi.shape != tuple(self.pad[::-1]):
            raise ValueError("psi must have shape %s" % (tuple(self.pad[::-1]),))
        _check(self.L.falcon_set_psi(self.h, ib, _p(psi)))

    def get_psi(self, ib):
        out = np.zeros(self.pad[::-1])
        _check(self.L.falcon_get_psi(self.h, ib, _p(out)))
        return out

    def init_synthetic(self, seed=12345, s_amp=1e-4, b_ratio=1e-3):
        _check(self.L.falcon_init_synthetic(self.h, seed, s_amp, b_ratio))

    # ---- stepping
    def set_precision(self, mode):
        _check(self.L.falcon_set_precision(self.h, {"fp64": 0, "fp32": 1}.get(mode, mode)))

    def set_ghost_pull(self, on):
        """FP32 step: stage/diffusion kernels read same-level ghost copies from
        the neighbours' interiors (on, default) or the ghost fill writes every
        ghost (off)."""
        _check(self.L.falcon_set_ghost_pull(self.h, 1 if on else 0))

    def set_tiles(self, precision, stage, diffusion, stage_wgs=0, diffusion_wgs=0):
        """Launch shapes of the stage and diffusion marches of one precision
        (0 FP64, 1 FP32 deviation): (theta, phi) tile edge, 8 or 16, and
        workgroups (0: one per tile; n > 0: n workgroups looping over the
        tiles).  Results do not depend on them."""
        _check(self.L.falcon_set_tiles(self.h, int(precision), int(stage), int(stage_wgs), int(diffusion),
                                       int(diffusion_wgs)))

    def get_tiles(self, precision):
        """(stage edge, stage workgroups, diffusion edge, diffusion workgroups)."""
        sh = (C.c_int32 * 4)()
        _check(self.L.falcon_get_tiles(self.h, int(precision), sh))
        return tuple(sh)

    SHAPES = ("16/tile", "16/persistent", "8/tile", "8/persistent")

    def autotune(self, reps=5):
        """Time the stage and diffusion kernels of the current precision in
        four launch shapes (tile edge 16 or 8, one workgroup per tile or the
        persistent grid) on the current state and keep the fastest of each.
        Returns {"stage": [edge, wgs], "diffusion": [edge, wgs],
        "ms": {kernel: {shape: ms}}}."""
        sh = (C.c_int32 * 4)()
        ms = (C.c_double * 8)()
        _check(self.L.falcon_autotune(self.h, int(reps), sh, ms))
        return {"stage": [sh[0], sh[1]], "diffusion": [sh[2], sh[3]],
                "ms": {"stage": dict(zip(self.SHAPES, ms[0:4])), "diffusion": dict(zip(self.SHAPES, ms[4:8]))}}

    def sendrecv_gc(self, which=0):
        _check(self.L.falcon_sendrecv_gc(self.h, which))

    def advance_rk4(self, cfl=0.5):
        dt = C.c_double()
        _check(self.L.falcon_advance_rk4(self.h, cfl, C.byref(dt)))
        return dt.value

    def advance(self, nsteps, cfl=0.5, graph=True):
        dt = C.c_double()
        _check(self.L.falcon_advance(self.h, nsteps, cfl, 1 if graph else 0, C.byref(dt)))
        return dt.value

    def compute_dt(self, cfl=0.5):
        dt = C.c_double()
        _check(self.L.falcon_compute_dt(self.h, cfl, C.byref(dt)))
        return dt.value

    def compute_dt_noalloc(self, cfl=0.5):
        """compute_dt through the library's static scratch (no device allocation)."""
        dt = C.c_double()
        _check(self.L.falcon_compute_dt_noalloc(self.h, cfl, C.byref(dt)))
        return dt.value

    def check_state(self, which=0):
        """ModCheck_primitive on the device: (NaN count, first NaN as (block,
        I, J, K padded 0-based, variable) or None); which 0 = state, 1 = stage buffer."""
        n = C.c_int64()
        w = np.zeros(5, np.int32)
        _check(self.L.falcon_check_state(self.h, int(which), C.byref(n), _p(w, C.c_int32)))
        return n.value, (tuple(int(x) for x in w) if n.value else None)

    def save_allcells(self, step, outdir, merged=False):
        """ModSave_AllCells output of the HBM state: outdir/AllCells_<step>/ with
        info.dat and Block_<label>.dat files (merged: merged.dat, the Tecplot
        file of merge_AllCells.py); returns that directory."""
        _check(self.L.falcon_save_allcells(self.h, int(step), os.fsencode(outdir), 1 if merged else 0))
        return os.path.join(outdir, "AllCells_%08d" % int(step))

    def set_fixed_dt(self, dt):
        """dt > 0: later steps use this dt instead of the CFL dt; 0 restores it."""
        _check(self.L.falcon_set_fixed_dt(self.h, float(dt)))

    def state_diff(self, ref, rel_tol=1e-6):
        """Interior difference to another Shell's state (same layout), per
        variable: max|a-b|, max|ref|, cells above rel_tol*max|ref| and the
        (block, i, j, k) of the largest difference."""
        md, mr = np.zeros(8), np.zeros(8)
        na, am = np.zeros(8, np.int64), np.zeros(8, np.int64)
        _check(self.L.falcon_state_diff(self.h, ref.h, float(rel_tol), _p(md), _p(mr), _p(na, C.c_int64),
                                        _p(am, C.c_int64)))
        where = []
        for x in am.tolist():
            if x < 0:
                where.append(None)
                continue
            k = x % self.nk; x //= self.nk
            j = x % self.nj; x //= self.nj
            i = x % self.ni; b = x // self.ni
            where.append((int(b), int(i), int(j), int(k)))
        return {"max_diff": md, "max_ref": mr, "n_above": na, "argmax": where}

    def rhs(self):
        out = np.zeros((self.nblocks, 8, self.nk, self.nj, self.ni))
        _check(self.L.falcon_rhs(self.h, _p(out)))
        return out

    def diffusion_rhs(self):
        out = np.zeros((self.nblocks, 8, self.nk, self.nj, self.ni))
        _check(self.L.falcon_diffusion_rhs(self.h, _p(out)))
        return out

    def profile_step(self, nsteps, cfl=0.5):
        out = np.zeros(7)
        _check(self.L.falcon_profile_step(self.h, nsteps, cfl, _p(out)))
        keys = ["dt", "boundary", "stage", "exchange_stage", "exchange_state", "diffusion", "step"]
        return dict(zip(keys, out.tolist()))

    def device_bytes(self):
        b = C.c_int64()
        _check(self.L.falcon_device_bytes(self.h, C.byref(b)))
        return b.value

    def synchronize(self):
        _check(self.L.falcon_synchronize(self.h))

    @property
    def interior_cells(self):
        return self.nblocks * self.ni * self.nj * self.nk


def format_es16(values):
    """Fortran ES16.8 fields of the values (the AllCells text edit), one string."""
    v = np.ascontiguousarray(np.atleast_1d(values), dtype=np.float64)
    buf = C.create_string_buffer(16 * v.size + 1)
    _check(load().falcon_format_es16(_p(v), int(v.size), buf))
    return buf.value.decode()


class Pfss:
    """The PFSS potential field (3_5/ModPFSS.f90) on a Shell's blocks (cubic,
    power-of-two edge): the Laplacian with the flux Br S of a magnetogram
    through the bottom faces and phi = 0 on the top face, solved by the
    MG-preconditioned flexible CG; B = -grad(phi).  Magnetograms are arrays
    (nbottom, n, n) in the C view of Fortran (j, k) order (j fastest), one per
    bottom block (`bottom`, block order); phi (nblocks, n, n, n) and B
    (nblocks, 3, n, n, n) per block in the C view of Fortran order."""

    def __init__(self, shell):
        self.s = shell
        self.L = shell.L
        _check(self.L.falcon_pfss_setup(shell.h))
        nb = C.c_int32()
        _check(self.L.falcon_pfss_bottom_blocks(shell.h, None, C.byref(nb)))
        blocks = np.zeros(nb.value, np.int32)
        _check(self.L.falcon_pfss_bottom_blocks(shell.h, _p(blocks, C.c_int32), C.byref(nb)))
        self.bottom = blocks
        n = shell.ni
        self.n = n
        self.shape = (shell.nblocks, n, n, n)
        self.mshape = (len(blocks), n, n)

    def areas(self):
        a = np.zeros(self.mshape)
        _check(self.L.falcon_pfss_bottom_areas(self.s.h, _p(a)))
        return a

    def dipole_magnetogram(self, xyz, q, l):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        l = np.ascontiguousarray(l, dtype=np.float64).reshape(-1, 3)
        br = np.zeros(self.mshape)
        _check(self.L.falcon_pfss_dipole_magnetogram(self.s.h, len(q), _p(xyz), _p(q), _p(l), _p(br)))
        return br

    def set_magnetogram(self, br):
        br = np.ascontiguousarray(br, dtype=np.float64).reshape(self.mshape)
        _check(self.L.falcon_pfss_set_magnetogram(self.s.h, _p(br)))

    def solve(self, tol=1e-10, max_iter=500):
        it, rel, sec = C.c_int32(), C.c_double(), C.c_double()
        _check(self.L.falcon_pfss_solve(self.s.h, tol, max_iter, C.byref(it), C.byref(rel), C.byref(sec)))
        return it.value, rel.value, sec.value

    def phi(self):
        out = np.zeros(self.shape)
        _check(self.L.falcon_pfss_get_phi(self.s.h, _p(out)))
        return out

    def field(self):
        out = np.zeros((self.shape[0], 3) + self.shape[1:])
        _check(self.L.falcon_pfss_field(self.s.h, _p(out)))
        return out

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape((1,) + self.shape)
        y = np.zeros((1,) + self.shape)
        _check(self.L.falcon_heat_apply(self.s.h, _p(x), _p(y)))
        return y[0]

    def vcycle(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape((1,) + self.shape)
        z = np.zeros((1,) + self.shape)
        _check(self.L.falcon_heat_vcycle(self.s.h, _p(r), _p(z)))
        return z[0]


class HeatSolver:
    """Implicit heat conduction on a Shell's blocks (cubic, power-of-two edge):
    backward Euler, MG-preconditioned flexible CG, `nsys` systems solved in the
    same launches.  Interior fields are numpy arrays of shape
    (nsys, nblocks, n, n, n) in the C view of Fortran order per block (i fastest)."""

    def __init__(self, shell, kappa0=1.0, r_ref_rsun=0.7, nsys=1):
        self.s = shell
        self.L = shell.L
        self.nsys = nsys
        _check(self.L.falcon_heat_setup(shell.h, kappa0, r_ref_rsun, nsys))
        lv, cells, dte = C.c_int32(), C.c_int64(), C.c_double()
        _check(self.L.falcon_heat_info(shell.h, C.byref(lv), C.byref(cells), C.byref(dte)))
        self.levels, self.cells, self.dt_explicit = lv.value, cells.value, dte.value
        n = shell.ni
        self.shape = (nsys, shell.nblocks, n, n, n)

    def set_dt(self, dt):
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.nsys,)))
        _check(self.L.falcon_heat_set_dt(self.s.h, _p(dt)))

    def set_T(self, T):
        T = np.ascontiguousarray(T, dtype=np.float64).reshape(self.shape)
        _check(self.L.falcon_heat_set_T(self.s.h, _p(T)))

    def seed_T(self, seed, sigma):
        _check(self.L.falcon_heat_seed_T(self.s.h, seed, sigma))

    def solve(self, tol=1e-10, max_iter=500, use_corrector=False):
        it = np.zeros(self.nsys, np.int32)
        rel = np.zeros(self.nsys)
        sec = C.c_double()
        _check(self.L.falcon_heat_solve(self.s.h, tol, max_iter, 1 if use_corrector else 0, _p(it, C.c_int32), _p(rel),
                                        C.byref(sec)))
        return it, rel, sec.value

    def get_T(self, sys=0):
        out = np.zeros(self.shape[1:])
        _check(self.L.falcon_heat_get_T(self.s.h, sys, _p(out)))
        return out

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.shape)
        y = np.zeros(self.shape)
        _check(self.L.falcon_heat_apply(self.s.h, _p(x), _p(y)))
        return y

    def vcycle(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.shape)
        z = np.zeros(self.shape)
        _check(self.L.falcon_heat_vcycle(self.s.h, _p(r), _p(z)))
        return z

    def train_corrector(self, npairs=2000, blocks_per_pair=16, dt_range=(10.0, 300.0), max_solves=50, epochs=10,
                        batch=128, lr=2e-3, seed=1):
        """Train the learned V-cycle correction on residual/error snapshots
        generated by solves of this solver's systems (blocks_per_pair block
        images each) and attach it; returns (loss, pairs, seconds)."""
        loss, n, sec = C.c_double(), C.c_int64(), C.c_double()
        _check(self.L.falcon_heat_train_corrector(self.s.h, npairs, blocks_per_pair, dt_range[0], dt_range[1], max_solves,
                                                  epochs, batch, lr, seed, C.byref(loss), C.byref(n), C.byref(sec)))
        return loss.value, n.value, sec.value

    def attach_corrector(self):
        _check(self.L.falcon_heat_attach_corrector(self.s.h))

    def set_graph(self, on):
        """Replay each FCG iteration as a captured hipGraph (default) or launch eagerly."""
        _check(self.L.falcon_heat_set_graph(self.s.h, 1 if on else 0))

    def set_learned_blockstep(self, on):
        """Learned correction step per 16^3 block (on, default) or per system (off)."""
        _check(self.L.falcon_heat_set_learned_blockstep(self.s.h, 1 if on else 0))

    def set_training_loss(self, h1_alpha=-1.0, weights=(1.0, 1.0, 1.0), scale_free=False):
        """Network training loss: h1_alpha < 0 mean squared error (default);
        h1_alpha >= 0 the gradient-weighted error h1_alpha |d|^2 + sum_a w_a
        |d_a d|^2, weights = (w_r, w_theta, w_phi); scale_free: the fraction
        of the target's norm that the best multiple of the output leaves."""
        _check(self.L.falcon_cnn_set_loss(self.s.h, float(h1_alpha), *[float(w) for w in weights],
                                          1 if scale_free else 0))

    def set_learned_iters(self, n):
        """Build the first n search directions of a learned solve with the learned
        correction and the rest with the plain V-cycle (0: all of them)."""
        _check(self.L.falcon_heat_set_learned_iters(self.s.h, int(n)))

    def set_coarse(self, tol, max_iter):
        """Stop of the V-cycle's coarsest-level CG (relative tol, iteration cap)."""
        _check(self.L.falcon_heat_set_coarse(self.s.h, float(tol), int(max_iter)))

    def set_learned_sweeps(self, pre0, post0):
        """Finest-level smoothing sweeps of the V-cycle used with the learned correction."""
        _check(self.L.falcon_heat_set_learned_sweeps(self.s.h, pre0, post0))


class Cnn:
    """The learned correction's network of a Shell's handle (forward, gradient,
    training on given images).  Images are float32 arrays (n, 16, 16, 16)."""
    NPARAM = 7809

    def __init__(self, shell, seed=None):
        self.s, self.L = shell, shell.L
        if seed is not None:
            _check(self.L.falcon_cnn_init(shell.h, seed))

    @property
    def params(self):
        p = np.zeros(self.NPARAM, np.float32)
        _check(self.L.falcon_cnn_get_params(self.s.h, _p(p, C.c_float)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.NPARAM
        _check(self.L.falcon_cnn_set_params(self.s.h, _p(p, C.c_float)))

    def set_loss(self, h1_alpha=-1.0, weights=(1.0, 1.0, 1.0), scale_free=False):
        """Training loss: h1_alpha < 0 mean squared error (default); h1_alpha >= 0
        the gradient-weighted error h1_alpha |d|^2 + sum_a w_a |d_a d|^2 per
        image, weights = (w_r, w_theta, w_phi); scale_free: the fraction of the
        target's norm that the best multiple of the output leaves."""
        _check(self.L.falcon_cnn_set_loss(self.s.h, float(h1_alpha), *[float(w) for w in weights],
                                          1 if scale_free else 0))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros_like(x)
        _check(self.L.falcon_cnn_forward(self.s.h, x.shape[0], _p(x, C.c_float), _p(out, C.c_float)))
        return out

    def gradient(self, x, gout):
        x = np.ascontiguousarray(x, dtype=np.float32)
        gout = np.ascontiguousarray(gout, dtype=np.float32)
        g = np.zeros(self.NPARAM)
        _check(self.L.falcon_cnn_gradient(self.s.h, x.shape[0], _p(x, C.c_float), _p(gout, C.c_float), _p(g)))
        return g

    def train(self, x, y, epochs=1, batch=64, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_double()
        _check(self.L.falcon_cnn_train(self.s.h, x.shape[0], _p(x, C.c_float), _p(y, C.c_float), epochs, batch, lr,
                                       seed, C.byref(loss)))
        return loss.value
