"""The learned V-cycle correction on the GPU against the CPU oracle.

Network (MFMA, BF16 operands, FP32 accumulation) against cnn_oracle.c run with
the same bfloat16 rounding points in FP64: outputs and parameter gradients
agree to the accumulation-order differences (bfloat16 re-rounding of a few
activations).  Learned FCG against a restatement built from the oracle's
operator, V-cycle and network; and against plain MG-FCG on solver-generated
training data.
"""
import numpy as np
import pytest

import pyoracle
from falcon_amd import Cnn, HeatSolver, Shell, configs

pytestmark = pytest.mark.gpu

NP = pyoracle.CNN_NPARAM
GROUPS = (("W1", 0, 432), ("b1", 432, 448), ("W2", 448, 7360), ("b2", 7360, 7376), ("W3", 7376, 7808),
          ("b3", 7808, 7809))


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    p[0:432] = rng.standard_normal(432) * np.sqrt(2 / 27) * scale
    p[432:448] = 0.05 * rng.standard_normal(16)
    p[448:7360] = rng.standard_normal(6912) * np.sqrt(2 / 432) * scale
    p[7360:7376] = 0.05 * rng.standard_normal(16)
    p[7376:7808] = rng.standard_normal(432) * np.sqrt(1 / 432) * scale
    p[7808] = 0.01
    return p


def _shell():
    return Shell(configs.cubic((16, 32, 96), 16))


def test_forward_matches_oracle():
    s = _shell()
    c = Cnn(s)
    p = _params(1)
    c.params = p
    x = np.random.default_rng(2).standard_normal((3, 16, 16, 16)).astype(np.float32)
    print("[cnn forward] 3 images", flush=True)
    gpu = c.forward(x)
    ref = pyoracle.CnnOracle().forward(p, x, bf16=True)[0]
    err = np.max(np.abs(gpu - ref)) / np.max(np.abs(ref))
    print("max rel err", err, flush=True)
    assert err < 2e-3


def test_gradient_matches_oracle():
    s = _shell()
    c = Cnn(s)
    p = _params(3)
    c.params = p
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 16, 16, 16)).astype(np.float32)
    gout = (1e-3 * rng.standard_normal((2, 16, 16, 16))).astype(np.float32)
    print("[cnn gradient] 2 images", flush=True)
    g = c.gradient(x, gout)
    o = pyoracle.CnnOracle()
    _, h1, h2 = o.forward(p, x, bf16=True)
    ref = o.backward(p, x, h1, h2, gout.astype(np.float64), bf16=True)
    for name, a, b in GROUPS:
        err = np.max(np.abs(g[a:b] - ref[a:b])) / np.max(np.abs(ref[a:b]))
        print(name, "max rel err", err, flush=True)
        assert err < 1e-2, name


def test_training_fits_a_teacher():
    """Student trained by the GPU kernels (forward, backward, Adam) on images
    labelled by a teacher network: the loss falls well below the labels' power."""
    s = _shell()
    c = Cnn(s, seed=7)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((64, 16, 16, 16)).astype(np.float32)
    y = pyoracle.CnnOracle().forward(_params(6), x, bf16=False)[0].astype(np.float32)
    power = float(np.mean(y.astype(np.float64) ** 2))
    print("[cnn train] 64 images, label power", power, flush=True)
    loss = c.train(x, y, epochs=60, batch=16, lr=3e-3, seed=1)
    print("final loss", loss, flush=True)
    assert loss < 0.2 * power


@pytest.mark.parametrize("scale_free", [False, True])
def test_gradient_weighted_loss_step(scale_free):
    """One Adam step from fresh optimiser state moves every parameter by
    -lr sign(g).  Under the gradient-weighted losses the GPU step's signs equal
    those of the oracle's backward of the loss gradient, with the loss written
    with axis differences and differentiated by torch autograd."""
    import torch
    s = _shell()
    c = Cnn(s, seed=3)
    p = _params(4, scale=0.5)
    c.params = p
    rng = np.random.default_rng(11)
    x = rng.standard_normal((4, 16, 16, 16)).astype(np.float32)
    y = (0.3 * rng.standard_normal((4, 16, 16, 16))).astype(np.float32)
    alpha, lr, w = 0.05, 1e-4, (3.0, 1.0, 0.5)
    c.set_loss(alpha, w, scale_free)
    print(f"[cnn h1 loss step] 4 images scale_free={scale_free}", flush=True)
    c.train(x, y, epochs=1, batch=4, lr=lr, seed=0)
    step = p.astype(np.float64) - c.params.astype(np.float64)
    o = pyoracle.CnnOracle()
    out, h1, h2 = o.forward(p, x, bf16=True)
    cv = torch.tensor(out.astype(np.float64), requires_grad=True)
    e = torch.tensor(y.astype(np.float64))

    def inner(u, v):         # per image: alpha <u,v> + sum_a w_a <diff_a u, diff_a v>
        s = alpha * (u * v).sum(dim=(1, 2, 3))
        for a in (1, 2, 3):
            s = s + w[a - 1] * (torch.diff(u, dim=a) * torch.diff(v, dim=a)).sum(dim=(1, 2, 3))
        return s

    if scale_free:
        loss = (1.0 - inner(cv, e) ** 2 / (inner(cv, cv) * inner(e, e))).mean()
    else:
        loss = inner(cv - e, cv - e).sum() / cv.numel()
    loss.backward()
    g = o.backward(p, x, h1, h2, cv.grad.numpy(), bf16=True)
    big = np.abs(g) > 0.05 * np.abs(g).max()
    agree = float(np.mean(np.sign(step[big]) == np.sign(g[big])))
    print("params compared", int(big.sum()), "sign agreement", agree, flush=True)
    assert big.sum() > 100 and agree > 0.98
    # Adam's first step is lr |g|/(|g| + 1e-8): lr up to the epsilon
    assert np.allclose(np.abs(step[big]), lr, rtol=0.1)


# ---------------------------------------------------------------- learned FCG

def _fcg_learned_restated(H, cnn_oracle, params, dt, Tn, tol, maxit, blockstep=True, nlearn=0):
    """FCG(1) with the learned correction, restated from the oracle's operator
    (apply), V-cycle and network: z = M r, r' = r - A z, u = r'/diag, s = rms(u),
    c = s net(u/s), z += a c with a = <c,r'>/<c,Ac> per block (blockstep) or
    per system, 0 unless <c,Ac> > 0.  With nlearn > 0 only directions
    0 .. nlearn-1 take the correction, later ones are the plain z = M r."""
    b = (H.apply(1.0, np.ones_like(Tn)) / dt) * Tn
    diag = H.diag(dt)
    x = np.zeros_like(b); r = b.copy(); bn = np.linalg.norm(b)

    def prec(r, k):
        z = H.vcycle(dt, r)
        if nlearn and k >= nlearn:
            return z
        rp = r - H.apply(dt, z)
        u = rp / diag
        sc = np.sqrt(np.mean(u * u))
        img = (u / sc).astype(np.float32).transpose(0, 3, 2, 1)          # Fortran (k,j,i) -> [i][j][k]
        y = cnn_oracle.forward(params, img, bf16=True)[0].transpose(0, 3, 2, 1)
        cc = sc * y
        Ac = H.apply(dt, cc)
        nb = cc.shape[0]
        num = (cc * rp).reshape(nb, -1).sum(axis=1)
        den = (cc * Ac).reshape(nb, -1).sum(axis=1)
        if not blockstep:
            num, den = np.full(nb, num.sum()), np.full(nb, den.sum())
        a = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
        return z + a[:, None, None, None] * cc

    z = prec(r, 0); p = z.copy(); it = 0
    while True:
        q = H.apply(dt, p); g = np.vdot(p, q); al = np.vdot(p, r) / g
        x += al * p; r -= al * q; it += 1
        rel = np.linalg.norm(r) / bn
        if rel <= tol or it >= maxit:
            return x, it, rel
        z = prec(r, it); be = np.vdot(z, q) / g; p = z - be * p


@pytest.mark.parametrize("blockstep,nlearn", [(True, 0), (False, 0), (True, 3)])
def test_learned_fcg_matches_oracle_restatement(blockstep, nlearn):
    """Any network gives a convergent FCG through the step a (per block or per
    system); the GPU path and the restatement agree in iterations (within one)
    and solution, also when the plain V-cycle takes over after nlearn
    directions."""
    lay = configs.cubic((16, 32, 96), 16)
    o = pyoracle.Oracle(lay)
    H = pyoracle.HeatOracle(o, 1.0, 0.7)
    s = Shell(lay)
    hs = HeatSolver(s, 1.0, 0.7, 1)
    c = Cnn(s)
    p = _params(8, scale=0.5)
    c.params = p
    hs.attach_corrector()
    hs.set_learned_blockstep(blockstep)
    hs.set_learned_iters(nlearn)
    dt = 100.0 * hs.dt_explicit
    T = 1e-3 * np.random.default_rng(9).standard_normal(hs.shape)
    hs.set_dt(dt)
    hs.set_T(T)
    print(f"[learned fcg vs restatement] 24 blocks blockstep={blockstep} nlearn={nlearn}", flush=True)
    it, rel, _ = hs.solve(1e-10, 100, use_corrector=True)
    x = hs.get_T(0)
    xo, ito, relo = _fcg_learned_restated(H, pyoracle.CnnOracle(), p, dt, T[0], 1e-10, 100, blockstep, nlearn)
    print("gpu", int(it[0]), float(rel[0]), "restated", ito, relo, flush=True)
    assert rel[0] <= 1e-10 and abs(int(it[0]) - ito) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8


def test_trained_correction_beats_plain_mg_fcg():
    """Train on pairs generated by the solver's own systems (dt 10-300x the
    explicit limit), then solve fresh systems: the learned FCG reaches 1e-10
    in fewer iterations than plain MG-FCG, with the same solution."""
    lay = configs.cubic((32, 64, 192), 16)
    s = Shell(lay)
    hs = HeatSolver(s, 1.0, 0.7, 4)
    loss, npairs, sec = hs.train_corrector(npairs=256, blocks_per_pair=4, dt_range=(10.0, 300.0), max_solves=20,
                                           epochs=8, batch=64, lr=2e-3, seed=3)
    print("[train corrector] pairs", npairs, "loss", loss, "seconds", sec, flush=True)
    assert npairs == 256 and np.isfinite(loss)
    dts = np.array([10.0, 30.0, 100.0, 300.0]) * hs.dt_explicit
    T = 1e-3 * np.random.default_rng(10).standard_normal(hs.shape)
    hs.set_dt(dts)
    hs.set_T(T)
    it0, rel0, sec0 = hs.solve(1e-10, 200, use_corrector=False)
    x0 = [hs.get_T(k) for k in range(4)]
    it1, rel1, sec1 = hs.solve(1e-10, 200, use_corrector=True)
    print("plain", it0.tolist(), sec0, "learned", it1.tolist(), sec1, flush=True)
    assert np.all(rel1 <= 1e-10) and np.all(rel0 <= 1e-10)
    assert np.all(it1 <= it0) and it1[2] < it0[2] and it1[3] < it0[3]
    for k in range(4):
        x1 = hs.get_T(k)
        assert np.linalg.norm(x1 - x0[k]) / np.linalg.norm(x0[k]) < 1e-7


def test_learned_fcg_on_the_locally_refined_shell_32_systems():
    """The 32x64x192-per-panel shell with seeded local refinement (20% of the
    blocks split once, 5% twice), 32 heat systems with dt spread over 10-300x
    the explicit limit, the correction trained on 2000 solver-generated pairs:
    every system's learned FCG reaches 1e-10 in no more iterations than plain
    MG-FCG, and the two solutions agree to 1e-7."""
    print("[learned fcg] layout=shell_amr_32x64x192_b16 systems=32 pairs=2000", flush=True)
    s = Shell(configs.SHELL_AMR_32x64x192_B16)
    nsys = 32
    hs = HeatSolver(s, 1.0, 0.7, nsys)
    loss, npairs, sec = hs.train_corrector(npairs=2000, blocks_per_pair=4, dt_range=(10.0, 300.0), max_solves=50,
                                           epochs=8, batch=128, lr=2e-3, seed=5)
    print("[train corrector] pairs", npairs, "loss", loss, "seconds", sec, flush=True)
    assert npairs == 2000 and np.isfinite(loss)
    dts = np.geomspace(10.0, 300.0, nsys) * hs.dt_explicit
    T = 1e-3 * np.random.default_rng(11).standard_normal(hs.shape)
    hs.set_dt(dts)
    hs.set_T(T)
    it0, rel0, sec0 = hs.solve(1e-10, 200, use_corrector=False)
    x0 = [hs.get_T(k) for k in range(nsys)]
    it1, rel1, sec1 = hs.solve(1e-10, 200, use_corrector=True)
    print("plain", it0.tolist(), f"{sec0 * 1e3:.2f} ms", "learned", it1.tolist(), f"{sec1 * 1e3:.2f} ms", flush=True)
    assert np.all(rel1 <= 1e-10) and np.all(rel0 <= 1e-10)
    assert np.all(it1 <= it0)
    for k in range(nsys):
        x1 = hs.get_T(k)
        assert np.linalg.norm(x1 - x0[k]) / np.linalg.norm(x0[k]) < 1e-7
    s.close()
