// Seeded synthetic initial state for the configured shells: white-noise entropy
// perturbation and a solenoidal magnetic field B = curl A of a smooth random
// vector potential, evaluated in global Cartesian coordinates so that both
// Yin-Yang panels see the same field.  No reference counterpart beyond the
// initialisation routines of 3_6/ModInitiation.f90:14-188, which set other
// fields (harmonic/random velocity, uniform B_phi).
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "layout.hpp"

namespace falcon {

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline double u01(uint64_t x) { return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0); }

struct Mode { double k[3], a[3], ph; };

// Yang Cartesian (x',y',z') = (-x, z, y) of Yin/global Cartesian; the map is an
// involution, matching ModYinYang_CoordConv (3_6/ModYinYang.f90:8-17).
static void to_global(int yin, const double p[3], double g[3]) {
    if (yin) { g[0] = p[0]; g[1] = p[1]; g[2] = p[2]; }
    else { g[0] = -p[0]; g[1] = p[2]; g[2] = p[1]; }
}

// Fills interior cells of every block: state is [block][Fortran (NI,NJ,NK,8)].
void synthetic_state(const Layout &lay, const std::vector<double> &bg, uint64_t seed, double s_amp, double b_ratio,
                     int nthreads, std::vector<double> &state) {
    const Tree &t = lay.tree;
    const int ni = t.ni, nj = t.nj, nk = t.nk, ng = t.ng;
    const int NI = ni + 2 * ng, NJ = nj + 2 * ng, NK = nk + 2 * ng;
    const size_t P = (size_t)NI * NJ * NK;
    const int nb = lay.nblocks();
    state.assign((size_t)nb * 8 * P, 0.0);
    const double R = t.r_range[1];
    const int nm = 24;
    std::vector<Mode> modes(nm);
    for (int m = 0; m < nm; m++) {
        uint64_t s = seed * 1000003ull + 77ull * m;
        double kmag = 2.0 * M_PI / (R * (0.05 + 0.25 * u01(s + 1)));
        double ct = 2.0 * u01(s + 2) - 1.0, phi = 2.0 * M_PI * u01(s + 3);
        double st = std::sqrt(1.0 - ct * ct);
        double kv[3] = {kmag * st * std::cos(phi), kmag * st * std::sin(phi), kmag * ct};
        double av[3] = {2.0 * u01(s + 4) - 1.0, 2.0 * u01(s + 5) - 1.0, 2.0 * u01(s + 6) - 1.0};
        Mode &M = modes[m];
        // B contribution of A = a sin(k.x + ph) is (k x a) cos(k.x + ph)
        M.k[0] = kv[0]; M.k[1] = kv[1]; M.k[2] = kv[2];
        M.a[0] = kv[1] * av[2] - kv[2] * av[1];
        M.a[1] = kv[2] * av[0] - kv[0] * av[2];
        M.a[2] = kv[0] * av[1] - kv[1] * av[0];
        M.ph = 2.0 * M_PI * u01(s + 7);
    }
    std::vector<double> maxratio(nb, 0.0);
    auto work = [&](int b) {
        const BlockGeom &g = lay.geom[b];
        double *S = state.data() + (size_t)b * 8 * P;
        const double *B = bg.data() + (size_t)b * 8 * NI;
        double mr = 0.0;
        for (int k = 1; k <= nk; k++)
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++) {
                    const int I = i + ng - 1, J = j + ng - 1, K = k + ng - 1;
                    const size_t q = (size_t)I + (size_t)NI * (J + (size_t)NJ * K);
                    const double r = g.xi[I], th = g.xj[J], ph = g.xk[K];
                    const double sth = std::sin(th), cth = std::cos(th), sph = std::sin(ph), cph = std::cos(ph);
                    double p[3] = {r * sth * cph, r * sth * sph, r * cth}, x[3];
                    to_global(g.yin, p, x);
                    double bgl[3] = {0, 0, 0};
                    for (const Mode &M : modes) {
                        double c = std::cos(M.k[0] * x[0] + M.k[1] * x[1] + M.k[2] * x[2] + M.ph);
                        bgl[0] += M.a[0] * c; bgl[1] += M.a[1] * c; bgl[2] += M.a[2] * c;
                    }
                    double bl[3];
                    to_global(g.yin, bgl, bl);   // same involution maps global components back
                    const double rhat[3] = {sth * cph, sth * sph, cth};
                    const double that[3] = {cth * cph, cth * sph, -sth};
                    const double phat[3] = {-sph, cph, 0.0};
                    const double Br = bl[0] * rhat[0] + bl[1] * rhat[1] + bl[2] * rhat[2];
                    const double Bt = bl[0] * that[0] + bl[1] * that[1] + bl[2] * that[2];
                    const double Bp = bl[0] * phat[0] + bl[1] * phat[1] + bl[2] * phat[2];
                    S[q + 4 * P] = Br; S[q + 5 * P] = Bt; S[q + 6 * P] = Bp;
                    const double g1 = B[I + 4 * NI], p0 = B[I + 2 * NI], rho0 = B[I + NI], T0 = B[I + 3 * NI];
                    const double cs = std::sqrt(g1 * p0 / rho0);
                    const double ratio = std::sqrt(Br * Br + Bt * Bt + Bp * Bp) / std::sqrt(rho0) / cs;
                    if (ratio > mr) mr = ratio;
                    // c_p = gamma/(gamma-1) * p0/(rho0 T0)
                    const double g3 = B[I + 5 * NI];
                    const double cp = g3 / (g3 - 1.0) * p0 / (rho0 * T0);
                    const uint64_t id = ((uint64_t)b << 32) ^ (uint64_t)((i * 131 + j) * 1031 + k);
                    S[q + 7 * P] = s_amp * cp * (2.0 * u01(seed ^ (id * 0x2545F4914F6CDD1Dull)) - 1.0);
                }
        maxratio[b] = mr;
    };
    std::vector<std::thread> th;
    int nt = std::max(1, std::min(nthreads, nb));
    for (int w = 0; w < nt; w++) th.emplace_back([&, w]() { for (int b = w; b < nb; b += nt) work(b); });
    for (auto &x : th) x.join();
    double mx = 0.0;
    for (double v : maxratio) mx = v > mx ? v : mx;
    const double scale = mx > 0.0 ? b_ratio / mx : 0.0;
    for (int b = 0; b < nb; b++) {
        double *S = state.data() + (size_t)b * 8 * P;
        for (size_t q = 0; q < P; q++) { S[q + 4 * P] *= scale; S[q + 5 * P] *= scale; S[q + 6 * P] *= scale; }
    }
}

}  // namespace falcon
