"""CPU restatement of the seeded synthetic initial state (test infrastructure).

The product generates the state on the device (falcon_amd/csrc/synthetic.cpp
draws the modes, kernels.hip k_synthetic / k_synthetic_scale evaluate them);
this is the same definition in numpy, evaluated on the block centres:
- rho1 = v = 0;
- s1 = s_amp * c_p * (2u - 1), u = splitmix64(seed ^ (id * 0x2545F4914F6CDD1D))
  as a 53-bit fraction, id = (block << 32) ^ ((i*131 + j)*1031 + k) with
  Fortran cell indices;
- B = sum over 24 modes of a cos(k.x + ph) at the cell's global Cartesian
  position (Yang through (x, y, z) -> (-x, z, y)), in the panel's spherical
  components, scaled so that max |B| / (sqrt(rho0) c_s) = b_ratio.
Only tests import it.
"""
import numpy as np

_M64 = (1 << 64) - 1


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def _u01(x):
    return float(_splitmix64(x) >> 11) * (1.0 / 9007199254740992.0)


def modes(r_hi, seed):
    """24 x (k[3], a[3], ph) for the outer radius r_hi (cm)."""
    out = []
    for m in range(24):
        s = (seed * 1000003 + 77 * m) & _M64
        kmag = 2.0 * np.pi / (r_hi * (0.05 + 0.25 * _u01(s + 1)))
        ct = 2.0 * _u01(s + 2) - 1.0
        phi = 2.0 * np.pi * _u01(s + 3)
        st = np.sqrt(1.0 - ct * ct)
        kv = np.array([kmag * st * np.cos(phi), kmag * st * np.sin(phi), kmag * ct])
        av = np.array([2.0 * _u01(s + 4) - 1.0, 2.0 * _u01(s + 5) - 1.0, 2.0 * _u01(s + 6) - 1.0])
        a = np.array([kv[1] * av[2] - kv[2] * av[1], kv[2] * av[0] - kv[0] * av[2], kv[0] * av[1] - kv[1] * av[0]])
        out.append((kv, a, 2.0 * np.pi * _u01(s + 7)))
    return out


def state(blocks, r_hi, seed, s_amp, b_ratio, ng=2):
    """blocks: list of (xi, xj, xk, yin, bg) with padded centres and the 8 x NI
    background rows (g, rho0, p0, T0, gamma1, gamma3, heat, xi).  Returns the
    interior states (8, nk, nj, ni) per block."""
    md = modes(r_hi, seed)
    out, mx = [], 0.0
    for b, (xi, xj, xk, yin, bg) in enumerate(blocks):
        r, th, ph = xi[ng:-ng], xj[ng:-ng], xk[ng:-ng]
        ni, nj, nk = r.size, th.size, ph.size
        K, J, I = np.meshgrid(ph, th, r, indexing="ij")
        sth, cth, sph, cph = np.sin(J), np.cos(J), np.sin(K), np.cos(K)
        p = [I * sth * cph, I * sth * sph, I * cth]
        x = p if yin else [-p[0], p[2], p[1]]
        g = [np.zeros_like(I) for _ in range(3)]
        for kv, a, phs in md:
            c = np.cos(kv[0] * x[0] + kv[1] * x[1] + kv[2] * x[2] + phs)
            for q in range(3):
                g[q] = g[q] + a[q] * c
        bl = g if yin else [-g[0], g[2], g[1]]
        Br = bl[0] * (sth * cph) + bl[1] * (sth * sph) + bl[2] * cth
        Bt = bl[0] * (cth * cph) + bl[1] * (cth * sph) + bl[2] * (-sth)
        Bp = bl[0] * (-sph) + bl[1] * cph
        rows = bg[:, ng:-ng]
        rho0, p0, T0, g1, g3 = rows[1], rows[2], rows[3], rows[4], rows[5]
        cp = g3 / (g3 - 1.0) * p0 / (rho0 * T0)
        cs = np.sqrt(g1 * p0 / rho0)
        ratio = np.sqrt(Br * Br + Bt * Bt + Bp * Bp) / np.sqrt(rho0)[None, None, :] / cs[None, None, :]
        mx = max(mx, float(ratio.max()))
        s1 = np.zeros((nk, nj, ni))
        for k in range(1, nk + 1):
            for j in range(1, nj + 1):
                for i in range(1, ni + 1):
                    cid = ((b << 32) ^ ((i * 131 + j) * 1031 + k)) & _M64
                    u = _u01(seed ^ ((cid * 0x2545F4914F6CDD1D) & _M64))
                    s1[k - 1, j - 1, i - 1] = s_amp * cp[i - 1] * (2.0 * u - 1.0)
        st = np.zeros((8, nk, nj, ni))
        st[4], st[5], st[6], st[7] = Br, Bt, Bp, s1
        out.append(st)
    scale = b_ratio / mx if mx > 0 else 0.0
    for st in out:
        st[4:7] *= scale
    return out
