"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The product package (falcon_amd) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
HARNESS = os.path.join(HERE, "_ref", "falcon_ref_harness")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_destroy.argtypes = [vp]
        L.orc_divide_all.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.orc_amr.argtypes = [vp, ip, C.c_int, dp]
        L.orc_refine.argtypes = [vp, C.c_int, ip]
        L.orc_finalize.argtypes = [vp]
        L.orc_nblocks.argtypes = [vp]
        L.orc_nleaf_yin.argtypes = [vp]
        L.orc_block_info.argtypes = [vp, C.c_int, dp, ip]
        L.orc_block_centres.argtypes = [vp, C.c_int, dp, dp, dp]
        L.orc_gc_iblocks.argtypes = [vp, C.c_int, ip]
        L.orc_ntargets.argtypes = [vp, C.c_int]
        L.orc_target.argtypes = [vp, C.c_int, C.c_int, ip, dp]
        L.orc_target_donor.argtypes = [vp, C.c_int, C.c_int]
        L.orc_nsources.argtypes = [vp, C.c_int]
        L.orc_source.argtypes = [vp, C.c_int, C.c_int, ip, ip, dp]
        L.orc_set_background.argtypes = [vp, C.c_int, dp]
        L.orc_set_state.argtypes = [vp, C.c_int, dp]
        L.orc_get_state.argtypes = [vp, C.c_int, dp]
        L.orc_exchange.argtypes = [vp, C.c_int]
        L.orc_rhs.argtypes = [vp, C.c_int, dp]
        L.orc_diffusion.argtypes = [vp, C.c_int, dp]
        L.orc_dt.argtypes = [vp, C.c_double]
        L.orc_dt.restype = C.c_double
        L.orc_step.argtypes = [vp, C.c_double]
        L.orc_step.restype = C.c_double
        L.orc_set_linear_rhs.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        L.orc_set_glm.argtypes = [vp, C.c_int]
        L.orc_set_psi.argtypes = [vp, C.c_int, dp]
        L.orc_get_psi.argtypes = [vp, C.c_int, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_heat_create.restype = vp
        L.orc_heat_create.argtypes = [vp, C.c_double, C.c_double]
        L.orc_heat_destroy.argtypes = [vp]
        L.orc_heat_dt_explicit.restype = C.c_double
        L.orc_heat_dt_explicit.argtypes = [vp]
        L.orc_heat_levels.argtypes = [vp]
        L.orc_heat_apply.argtypes = [vp, C.c_double, dp, dp]
        L.orc_heat_vcycle.argtypes = [vp, C.c_double, dp, dp]
        L.orc_heat_diag.argtypes = [vp, C.c_double, dp]
        L.orc_heat_solve.argtypes = [vp, C.c_double, dp, C.c_double, C.c_int, dp, dp]
        L.orc_heat_set_coarse.argtypes = [vp, C.c_double, C.c_int]
        L.orc_heat_prolong.argtypes = [vp, C.c_int, dp, dp]
        L.orc_heat_create_opt.restype = vp
        L.orc_heat_create_opt.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_heat_solve_b.argtypes = [vp, C.c_double, dp, C.c_double, C.c_int, dp, dp]
        L.orc_pfss_field.argtypes = [vp, dp, dp, dp]
        fp = C.POINTER(C.c_float)
        L.cnn_round_bf16.restype = C.c_float
        L.cnn_round_bf16.argtypes = [C.c_float]
        L.cnn_forward.argtypes = [fp, C.c_int, fp, dp, dp, dp, C.c_int]
        L.cnn_backward.argtypes = [fp, C.c_int, fp, dp, dp, dp, dp, C.c_int]
        L.orc_coordconv.argtypes = [dp, dp]
        L.orc_vecconv_coeffs.argtypes = [dp, dp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """One Yin-Yang layout in the oracle: tree, blocks, tables, state."""

    def __init__(self, layout):
        L = lib()
        self.L = L
        self.layout = layout
        ni, nj, nk, ng = layout["ni"], layout["nj"], layout["nk"], layout["ng"]
        self.ni, self.nj, self.nk, self.ng = ni, nj, nk, ng
        self.pad = (ni + 2 * ng, nj + 2 * ng, nk + 2 * ng)
        self.h = L.orc_create(ni, nj, nk, ng, layout["r_rsun"][0], layout["r_rsun"][1])
        for op in layout.get("ops", []):
            if op[0] == "split":
                L.orc_divide_all(self.h, *op[1])
            elif op[0] == "refine":
                ids = np.asarray(op[1], dtype=np.int32)
                if L.orc_refine(self.h, len(ids), _p(ids, C.c_int32)) != 0:
                    raise ValueError("refine: leaf ordinal out of range")
            else:
                _, flags, glob, rng = op
                f = np.asarray(flags, dtype=np.int32)
                r = np.asarray(rng, dtype=np.float64)
                L.orc_amr(self.h, _p(f, C.c_int32), int(glob), _p(r, C.c_double))
        self.nblocks = L.orc_finalize(self.h)
        self.nleaf_yin = L.orc_nleaf_yin(self.h)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def block_info(self, ib):
        rng = np.zeros(6)
        fl = np.zeros(3, dtype=np.int32)
        self.L.orc_block_info(self.h, ib, _p(rng, C.c_double), _p(fl, C.c_int32))
        return rng, fl

    def centres(self, ib):
        xi, xj, xk = (np.zeros(n) for n in self.pad)
        self.L.orc_block_centres(self.h, ib, _p(xi, C.c_double), _p(xj, C.c_double), _p(xk, C.c_double))
        return xi, xj, xk

    def gc_iblocks(self, ib):
        out = np.zeros(self.pad[::-1], dtype=np.int32)  # C view of Fortran (i,j,k): [k][j][i]
        self.L.orc_gc_iblocks(self.h, ib, _p(out, C.c_int32))
        return out

    def targets(self, ib):
        res = []
        for it in range(self.L.orc_ntargets(self.h, ib)):
            n = self.L.orc_target(self.h, ib, it, None, None)
            ijk = np.zeros((3, n), dtype=np.int32)
            x = np.zeros((3, n))
            self.L.orc_target(self.h, ib, it, _p(ijk, C.c_int32), _p(x, C.c_double))
            res.append((self.L.orc_target_donor(self.h, ib, it), ijk, x))
        return res

    def sources(self, ib):
        res = []
        for s in range(self.L.orc_nsources(self.h, ib)):
            rb = np.zeros(1, dtype=np.int32)
            n = self.L.orc_source(self.h, ib, s, _p(rb, C.c_int32), None, None)
            ijk = np.zeros((3, n), dtype=np.int32)
            x = np.zeros((3, n))
            self.L.orc_source(self.h, ib, s, _p(rb, C.c_int32), _p(ijk, C.c_int32), _p(x, C.c_double))
            res.append((int(rb[0]), ijk, x))
        return res

    # state arrays are numpy arrays shaped (8, NK, NJ, NI): the C view of Fortran (i,j,k,v)
    def set_background(self, ib, bg):
        bg = np.ascontiguousarray(bg, dtype=np.float64)  # (8, NI)
        self.L.orc_set_background(self.h, ib, _p(bg, C.c_double))

    def set_state(self, ib, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self.L.orc_set_state(self.h, ib, _p(u, C.c_double))

    def get_state(self, ib):
        out = np.zeros((8,) + self.pad[::-1])
        self.L.orc_get_state(self.h, ib, _p(out, C.c_double))
        return out

    def set_glm(self, on):
        """GLM divergence cleaning (3_6/ModDivB.f90): psi becomes a ninth
        stepped variable."""
        self.L.orc_set_glm(self.h, 1 if on else 0)

    def set_psi(self, ib, psi):
        """padded psi of block ib, (NK, NJ, NI); written to the state and the
        stage buffer, ghosts included"""
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        assert psi.shape == tuple(self.pad[::-1])
        self.L.orc_set_psi(self.h, ib, _p(psi, C.c_double))

    def get_psi(self, ib):
        out = np.zeros(self.pad[::-1])
        self.L.orc_get_psi(self.h, ib, _p(out, C.c_double))
        return out

    def exchange(self, which=0):
        self.L.orc_exchange(self.h, which)

    def rhs(self, ib):
        out = np.zeros((8, self.nk, self.nj, self.ni))
        self.L.orc_rhs(self.h, ib, _p(out, C.c_double))
        return out

    def diffusion(self, ib):
        out = np.zeros((8, self.nk, self.nj, self.ni))
        self.L.orc_diffusion(self.h, ib, _p(out, C.c_double))
        return out

    def dt(self, cfl):
        return self.L.orc_dt(self.h, cfl)

    def step(self, cfl):
        return self.L.orc_step(self.h, cfl)

    def set_linear_rhs(self, on, lam=0.0, mu=0.0):
        """Pin hook: the step's stage right-hand side becomes lam * (stage input)
        and the diffusion's mu * U (see orc_set_linear_rhs)."""
        self.L.orc_set_linear_rhs(self.h, 1 if on else 0, float(lam), float(mu))


class HeatOracle:
    """Heat operator and MG-preconditioned FCG of the oracle on an Oracle layout
    (cubic blocks).  Interior arrays are (nblocks, n, n, n) in C view of Fortran
    (i fastest), i.e. numpy shape (nblocks, nk, nj, ni)."""

    def __init__(self, orc, kappa0, r_ref_rsun, kappa_power=-3, top_dirichlet=False):
        self.L = lib()
        self.o = orc
        self.h = self.L.orc_heat_create_opt(orc.h, kappa0, r_ref_rsun * float(np.float32(6.955e10)), int(kappa_power),
                                            int(bool(top_dirichlet)))
        self.n = orc.ni
        self.shape = (orc.nblocks, orc.nk, orc.nj, orc.ni)

    def __del__(self):
        try:
            self.L.orc_heat_destroy(self.h)
        except Exception:
            pass

    @property
    def dt_explicit(self):
        return self.L.orc_heat_dt_explicit(self.h)

    @property
    def levels(self):
        return self.L.orc_heat_levels(self.h)

    def apply(self, dt, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.shape)
        self.L.orc_heat_apply(self.h, dt, _p(x, C.c_double), _p(y, C.c_double))
        return y

    def prolong(self, level, zc, zf):
        """zf + P zc from level+1 (edge n/2^(level+1)) to level (test hook)."""
        zc = np.ascontiguousarray(zc, dtype=np.float64)
        out = np.array(zf, dtype=np.float64, copy=True, order="C")
        self.L.orc_heat_prolong(self.h, int(level), _p(zc, C.c_double), _p(out, C.c_double))
        return out

    def set_coarse(self, tol, max_iter):
        """Coarsest-level CG of the V-cycle: relative stop tol, iteration cap."""
        self.L.orc_heat_set_coarse(self.h, float(tol), int(max_iter))

    def vcycle(self, dt, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.shape)
        self.L.orc_heat_vcycle(self.h, dt, _p(r, C.c_double), _p(z, C.c_double))
        return z

    def diag(self, dt):
        d = np.zeros(self.shape)
        self.L.orc_heat_diag(self.h, dt, _p(d, C.c_double))
        return d

    def solve(self, dt, Tn, tol=1e-10, max_iter=500):
        Tn = np.ascontiguousarray(Tn, dtype=np.float64)
        x = np.zeros(self.shape)
        rel = np.zeros(1)
        it = self.L.orc_heat_solve(self.h, dt, _p(Tn, C.c_double), tol, max_iter, _p(x, C.c_double), _p(rel, C.c_double))
        return x, it, float(rel[0])

    def solve_b(self, dt, b, tol=1e-10, max_iter=500):
        """A x = b from x = 0 (dt = inf: no V/dt term)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.shape)
        rel = np.zeros(1)
        it = self.L.orc_heat_solve_b(self.h, dt, _p(b, C.c_double), tol, max_iter, _p(x, C.c_double), _p(rel, C.c_double))
        return x, it, float(rel[0])


class PfssOracle(HeatOracle):
    """The PFSS potential field of 3_5/ModPFSS.f90 as the oracle restates it:
    the Laplacian (uniform kappa, no V/dt term), insulated bottom faces whose
    flux b = Br S is the right-hand side, the value 0 on the top face (the
    source surface); B = -grad(phi), second order.  br: (nbottom, nk, nj), the
    bottom blocks in block order."""

    def __init__(self, orc):
        super().__init__(orc, 1.0, 1.0, kappa_power=0, top_dirichlet=True)
        self.bottom = [ib for ib in range(orc.nblocks) if orc.block_info(ib)[1][2]]

    def rhs(self, br, area):
        """b = Br S on the first radial cell of every bottom block (area: the
        bottom faces' areas, same shape as br)."""
        b = np.zeros(self.shape)
        for q, ib in enumerate(self.bottom):
            b[ib, :, :, 0] = br[q] * area[q]
        return b

    def solve_br(self, br, area, tol=1e-10, max_iter=500):
        return self.solve_b(float("inf"), self.rhs(br, area), tol, max_iter)

    def field(self, phi, br):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        br = np.ascontiguousarray(br, dtype=np.float64)
        B = np.zeros((self.o.nblocks, 3) + self.shape[1:])
        self.L.orc_pfss_field(self.h, _p(phi, C.c_double), _p(br, C.c_double), _p(B, C.c_double))
        return B


CNN_NPARAM = 7809


class CnnOracle:
    """The 1->16->16->1 3^3 network of the learned V-cycle correction (FP64).
    Images are float arrays (nimg, 16, 16, 16) indexed [i][j][k]."""

    def __init__(self):
        self.L = lib()

    def forward(self, params, x, bf16=True):
        params = np.ascontiguousarray(params, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        h1 = np.zeros((n, 4096, 16)); h2 = np.zeros((n, 4096, 16)); out = np.zeros((n, 16, 16, 16))
        self.L.cnn_forward(_p(params, C.c_float), n, _p(x, C.c_float), _p(h1, C.c_double), _p(h2, C.c_double),
                           _p(out, C.c_double), int(bf16))
        return out, h1, h2

    def backward(self, params, x, h1, h2, gout, bf16=True):
        params = np.ascontiguousarray(params, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        gout = np.ascontiguousarray(gout, dtype=np.float64)
        g = np.zeros(CNN_NPARAM)
        self.L.cnn_backward(_p(params, C.c_float), x.shape[0], _p(x, C.c_float), _p(h1, C.c_double), _p(h2, C.c_double),
                            _p(gout, C.c_double), _p(g, C.c_double), int(bf16))
        return g

    def round_bf16(self, a):
        f = np.vectorize(lambda v: self.L.cnn_round_bf16(float(v)))
        return f(np.asarray(a, dtype=np.float32)).astype(np.float32)


def set_threads(n):
    lib().orc_set_threads(int(n))
