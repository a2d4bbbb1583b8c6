#include "output.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace falcon {

void format_es16(double x, char *out) {
    char buf[40];
    if (std::isnan(x)) {
        std::snprintf(buf, sizeof(buf), "%16s", "NaN");
    } else if (std::isinf(x)) {
        std::snprintf(buf, sizeof(buf), "%16s", x > 0 ? "Infinity" : "-Infinity");
    } else {
        // printf's %E gives the correctly rounded 9 significant digits and at
        // least two exponent digits; Fortran's ESw.d drops the E when the
        // exponent needs three
        int n = std::snprintf(buf, sizeof(buf), "%.8E", x);
        char *e = std::strchr(buf, 'E');
        if (e && std::strlen(e + 2) == 3) {
            std::memmove(e, e + 1, std::strlen(e + 1) + 1);
            n--;
        }
        char tmp[40];
        std::snprintf(tmp, sizeof(tmp), "%16s", buf);
        std::memcpy(buf, tmp, 17);
        (void)n;
    }
    std::memcpy(out, buf, 16);
}

std::string allcells_info(const AllCellsSpec &s) {
    std::string o;
    o += "nBlocks = " + std::to_string(s.nblocks) + "\n";
    o += "ni = " + std::to_string(s.ni) + "\n";
    o += "nj = " + std::to_string(s.nj) + "\n";
    o += "nk = " + std::to_string(s.nk) + "\n";
    o += "nVarsTotal = " + std::to_string(s.nvar_total) + "\n";
    o += s.variables + "\n";
    return o;
}

void allcells_block_text(const AllCellsSpec &s, const double *vals, std::string &out) {
    const size_t nc = (size_t)(s.ni + 1) * (s.nj + 1) * (s.nk + 1);
    const size_t line = (size_t)s.nvar_total * 16 + 1;
    const size_t o0 = out.size();
    out.resize(o0 + nc * line);
    char *p = &out[o0];
    for (size_t c = 0; c < nc; c++) {
        for (int v = 0; v < s.nvar_total; v++) format_es16(vals[c * s.nvar_total + v], p + v * 16);
        p[s.nvar_total * 16] = '\n';
        p += line;
    }
}

void allcells_geometry(const Layout &lay, std::vector<double> &geo, std::vector<int32_t> &yin) {
    const Tree &t = lay.tree;
    const int ni = t.ni, nj = t.nj, nk = t.nk, nb = lay.nblocks();
    const size_t G = (size_t)(ni + 1) + 2 * (nj + 1) + 2 * (nk + 1);
    geo.assign(G * nb, 0.0);
    yin.assign(nb, 0);
    auto face = [](const double rng[2], int n, int f) {
        return (rng[0] * (double)(n + 1 - f) + rng[1] * (double)(f - 1)) / (double)n;
    };
    for (int b = 0; b < nb; b++) {
        const BlockGeom &g = lay.geom[b];
        double *o = geo.data() + G * b;
        yin[b] = g.yin;
        for (int i = 1; i <= ni + 1; i++) o[i - 1] = face(g.rng[0], ni, i);
        double *st = o + ni + 1, *ct = st + nj + 1, *sp = ct + nj + 1, *cp = sp + nk + 1;
        // sth = sin(theta); cth = cos(theta) (3_6/ModSavePlot.f90:601-602): one
        // sincos call per angle, as the reference's build evaluates the pair
        for (int j = 1; j <= nj + 1; j++) ::sincos(face(g.rng[1], nj, j), &st[j - 1], &ct[j - 1]);
        for (int k = 1; k <= nk + 1; k++) ::sincos(face(g.rng[2], nk, k), &sp[k - 1], &cp[k - 1]);
    }
}

void allcells_connectivity(const AllCellsSpec &s, int b0, int b1, std::string &out) {
    const long long sj = s.ni + 1, sk = (long long)(s.ni + 1) * (s.nj + 1);
    const long long per = sk * (s.nk + 1);
    char buf[200];
    for (int b = b0; b < b1; b++) {
        const long long off = (long long)b * per;
        for (int k = 1; k <= s.nk; k++)
            for (int j = 1; j <= s.nj; j++)
                for (int i = 1; i <= s.ni; i++) {
                    const long long c1 = off + (k - 1) * sk + (j - 1) * sj + i;
                    const long long c2 = c1 + 1, c3 = c1 + sj + 1, c4 = c1 + sj;
                    const int n = std::snprintf(buf, sizeof(buf), "%lld %lld %lld %lld %lld %lld %lld %lld\n", c1, c2,
                                                c3, c4, c1 + sk, c2 + sk, c3 + sk, c4 + sk);
                    out.append(buf, (size_t)n);
                }
    }
}

}  // namespace falcon
