// Seeded synthetic initial state for the configured shells: white-noise entropy
// perturbation and a solenoidal magnetic field B = curl A of a smooth random
// vector potential, evaluated in global Cartesian coordinates so that both
// Yin-Yang panels see the same field.  No reference counterpart beyond the
// initialisation routines of 3_6/ModInitiation.f90:14-188, which set other
// fields (harmonic/random velocity, uniform B_phi).
//
// The host draws the potential's Fourier modes; the device evaluates the
// field and the noise on every cell (k_synthetic).
#include <cmath>
#include <cstdint>
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

// 24 modes, 7 numbers each: k (3), a = k x (random amplitude) (3), phase.
// Wavelengths are 0.05-0.3 of the outer radius, directions and amplitudes
// uniform.  The B contribution of A = amp sin(k.x + ph) is a cos(k.x + ph).
void synthetic_modes(const Tree &t, uint64_t seed, std::vector<double> &modes) {
    const double R = t.r_range[1];
    const int nm = 24;
    modes.assign((size_t)7 * nm, 0.0);
    for (int m = 0; m < nm; m++) {
        uint64_t s = seed * 1000003ull + 77ull * m;
        double kmag = 2.0 * M_PI / (R * (0.05 + 0.25 * u01(s + 1)));
        double ct = 2.0 * u01(s + 2) - 1.0, phi = 2.0 * M_PI * u01(s + 3);
        double st = std::sqrt(1.0 - ct * ct);
        double kv[3] = {kmag * st * std::cos(phi), kmag * st * std::sin(phi), kmag * ct};
        double av[3] = {2.0 * u01(s + 4) - 1.0, 2.0 * u01(s + 5) - 1.0, 2.0 * u01(s + 6) - 1.0};
        double *M = modes.data() + 7 * m;
        M[0] = kv[0]; M[1] = kv[1]; M[2] = kv[2];
        M[3] = kv[1] * av[2] - kv[2] * av[1];
        M[4] = kv[2] * av[0] - kv[0] * av[2];
        M[5] = kv[0] * av[1] - kv[1] * av[0];
        M[6] = 2.0 * M_PI * u01(s + 7);
    }
}

}  // namespace falcon
