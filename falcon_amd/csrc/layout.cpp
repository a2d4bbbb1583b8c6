// Host-side layout builder (see layout.hpp).  Compiled with -ffp-contract=off:
// every floating-point expression below is evaluated in the order the cited
// reference line evaluates it, so tables are bit-identical to the reference's.
#include "layout.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>

namespace falcon {

// 3_6/ModConst.f90:2,5 are default-real (float) parameters.
static const float kDpiF = 3.1415926535f;
static const float kRsunF = 6.955e10f;

// sin and cos of one angle come from one sincos call, as the reference's
// build evaluates the pairs it writes (LLVM fuses them into glibc's sincos,
// whose cos can differ from cos by one ulp)
void coordconv(const double in[3], double out[3]) {
    double st, ct, sp, cp;
    ::sincos(in[1], &st, &ct);
    ::sincos(in[2], &sp, &cp);
    out[0] = in[0];
    out[1] = std::acos(st * sp);
    double a = std::acos(-((st * cp) / std::sqrt((st * st) * (cp * cp) + ct * ct)));
    out[2] = std::copysign(std::fabs(a), ct);  // SIGN(a, cos t)
}

void sin_and_cot(double t, double &s, double &cot) {
    s = std::sin(t);
    double ss, cc;
    ::sincos(t, &ss, &cc);
    cot = cc / ss;
}

void Tree::init(int ni_, int nj_, int nk_, int ng_, double r_lo_cgs, double r_hi_cgs) {
    ni = ni_; nj = nj_; nk = nk_; ng = ng_;
    r_range[0] = r_lo_cgs; r_range[1] = r_hi_cgs;
    nodes.clear();
    Node root;
    root.rng[0][0] = r_lo_cgs; root.rng[0][1] = r_hi_cgs;
    root.rng[1][0] = (double)(kDpiF * 0.25f); root.rng[1][1] = (double)(kDpiF * 0.75f);
    root.rng[2][0] = (double)(-kDpiF * 0.75f); root.rng[2][1] = (double)(kDpiF * 0.75f);
    root.yin = 1;
    nodes.push_back(root);
    root.yin = 0;
    nodes.push_back(root);
    nleaf = nleaf_yin = 0;
    leaf_node.clear();
}

// Split rule of YinYangTree_Divide (3_6/ModYinYangTree.f90:245-255): r by a
// geometric rule with default-real exponents, theta/phi linearly.  N per axis is
// general; N in {1,2} is the reference's own case.
void Tree::divide(int node, const int N[3]) {
    if (!nodes[node].leaf) return;
    Node parent = nodes[node];
    int first = (int)nodes.size();
    for (int ir = 1; ir <= N[0]; ir++)
        for (int it = 1; it <= N[1]; it++)
            for (int ip = 1; ip <= N[2]; ip++) {
                Node c;
                c.yin = parent.yin;
                float ea0 = ((float)(N[0] - ir) + 1.0f) / (float)N[0];
                float ea1 = ((float)(N[0] - ir) + 0.0f) / (float)N[0];
                float eb0 = ((float)(-ir) + 1.0f) / (float)N[0];
                float eb1 = ((float)(-ir) + 0.0f) / (float)N[0];
                c.rng[0][0] = std::pow(parent.rng[0][0], (double)ea0) / std::pow(parent.rng[0][1], (double)eb0);
                c.rng[0][1] = std::pow(parent.rng[0][0], (double)ea1) / std::pow(parent.rng[0][1], (double)eb1);
                const int n2[2] = {N[1], N[2]};
                const int idx[2] = {it, ip};
                for (int a = 0; a < 2; a++) {
                    float fa0 = (float)(n2[a] - idx[a]) + 1.0f, fa1 = (float)(n2[a] - idx[a]) + 0.0f;
                    float fb0 = (float)(-idx[a]) + 1.0f, fb1 = (float)(-idx[a]) + 0.0f;
                    double lo = parent.rng[a + 1][0], hi = parent.rng[a + 1][1], n = (double)n2[a];
                    c.rng[a + 1][0] = (lo * (double)fa0) / n - (hi * (double)fb0) / n;
                    c.rng[a + 1][1] = (lo * (double)fa1) / n - (hi * (double)fb1) / n;
                }
                nodes.push_back(c);
            }
    Node &p = nodes[node];
    p.first_child = first;
    p.N[0] = N[0]; p.N[1] = N[1]; p.N[2] = N[2];
    p.leaf = 0;
}

void Tree::divide_all(int nr, int nt, int np) {
    const int N[3] = {nr, nt, np};
    std::vector<int> stack = {1, 0};
    std::vector<int> leaves;
    // collect current leaves first (depth-first, children in order), then divide
    while (!stack.empty()) {
        int n = stack.back(); stack.pop_back();
        if (nodes[n].leaf) { leaves.push_back(n); continue; }
        int nc = nodes[n].N[0] * nodes[n].N[1] * nodes[n].N[2];
        for (int c = nc - 1; c >= 0; c--) stack.push_back(nodes[n].first_child + c);
    }
    for (int n : leaves) divide(n, N);
}

// ModMath_IfLinesInterSect (3_6/ModMath.f90:5-16)
static bool lines_intersect(const double a[2], const double b[2]) {
    return (a[0] - b[0]) * (a[0] - b[1]) < 0.0 || (a[1] - b[0]) * (a[1] - b[1]) < 0.0 ||
           (b[0] - a[0]) * (b[0] - a[1]) < 0.0 || (b[1] - a[0]) * (b[1] - a[1]) < 0.0;
}

// ModAMR_Divide_OneBranch (3_6/ModAMR.f90:65-99) over both panels
void Tree::amr(const int flags[3], int global, const double rng_rsun[6]) {
    const int N[3] = {flags[0] ? 2 : 1, flags[1] ? 2 : 1, flags[2] ? 2 : 1};
    double rr[2] = {rng_rsun[0] * (double)kRsunF, rng_rsun[3] * (double)kRsunF};
    double tr[2] = {rng_rsun[1], rng_rsun[4]};
    double pr[2] = {rng_rsun[2], rng_rsun[5]};
    std::vector<int> leaves;
    std::vector<int> stack = {1, 0};
    while (!stack.empty()) {
        int n = stack.back(); stack.pop_back();
        if (nodes[n].leaf) { leaves.push_back(n); continue; }
        int nc = nodes[n].N[0] * nodes[n].N[1] * nodes[n].N[2];
        for (int c = nc - 1; c >= 0; c--) stack.push_back(nodes[n].first_child + c);
    }
    for (int n : leaves) {
        const Node &nd = nodes[n];
        bool go;
        if (global) go = lines_intersect(nd.rng[0], rr);
        else go = nd.yin && lines_intersect(nd.rng[0], rr) && lines_intersect(nd.rng[1], tr) && lines_intersect(nd.rng[2], pr);
        if (go) divide(n, N);
    }
}

// Local refinement: split the listed leaves (0-based ordinals in the current
// depth-first order, Yin then Yang) into 2 x 2 x 2 children.
void Tree::refine(const std::vector<int> &ordinals) {
    std::vector<int> leaves;
    std::vector<int> stack = {1, 0};
    while (!stack.empty()) {
        int n = stack.back(); stack.pop_back();
        if (nodes[n].leaf) { leaves.push_back(n); continue; }
        int nc = nodes[n].N[0] * nodes[n].N[1] * nodes[n].N[2];
        for (int c = nc - 1; c >= 0; c--) stack.push_back(nodes[n].first_child + c);
    }
    const int N[3] = {2, 2, 2};
    for (int o : ordinals) {
        if (o < 0 || o >= (int)leaves.size()) throw std::runtime_error("refine: leaf ordinal out of range");
        divide(leaves[o], N);
    }
}

// YinYangTree_LabelNodes (3_6/ModYinYangTree.f90:402-420): depth-first, Yin then Yang
void Tree::label() {
    nleaf = 0;
    leaf_node.clear();
    for (int root = 0; root < 2; root++) {
        std::vector<int> stack = {root};
        while (!stack.empty()) {
            int n = stack.back(); stack.pop_back();
            if (nodes[n].leaf) {
                nodes[n].leaf_id = ++nleaf;
                leaf_node.push_back(n);
                continue;
            }
            int nc = nodes[n].N[0] * nodes[n].N[1] * nodes[n].N[2];
            for (int c = nc - 1; c >= 0; c--) stack.push_back(nodes[n].first_child + c);
        }
        if (root == 0) nleaf_yin = nleaf;
    }
}

// YinYangTree_IfInsideNode (3_6/ModYinYangTree.f90:424-445); the point in the
// other panel's frame is computed once by the caller (CoordConv is deterministic).
bool Tree::inside(int node, const double rtp[3], const double rtp_other[3], int yin) const {
    const Node &n = nodes[node];
    const double *h = (yin == n.yin) ? rtp : rtp_other;
    return (h[0] - n.rng[0][0]) * (h[0] - n.rng[0][1]) <= 0.0 &&
           (h[1] - n.rng[1][0]) * (h[1] - n.rng[1][1]) <= 0.0 &&
           (h[2] - n.rng[2][0]) * (h[2] - n.rng[2][1]) <= 0.0;
}

// YinYangTree_FindNode_OneBranch (3_6/ModYinYangTree.f90:468-505): the first
// child, in child order, that contains the point and whose subtree resolves it.
// For an N-ary node only children whose per-axis intervals contain the point
// can pass the inside test, so they are enumerated per axis instead of scanning
// all N0*N1*N2 children; the visiting order is unchanged.
int Tree::find_branch(int node, const double rtp[3], const double rtp_other[3], int yin) const {
    if (!inside(node, rtp, rtp_other, yin)) return -777;
    const Node &n = nodes[node];
    if (n.leaf) return n.leaf_id;
    const int N0 = n.N[0], N1 = n.N[1], N2 = n.N[2];
    const int nc = N0 * N1 * N2;
    if (nc <= 8) {
        for (int c = 0; c < nc; c++) {
            int ch = n.first_child + c;
            int r = -777;
            if (inside(ch, rtp, rtp_other, yin)) r = find_branch(ch, rtp, rtp_other, yin);
            if (r != -777) return r;
        }
        return -777;
    }
    const double *h = (yin == n.yin) ? rtp : rtp_other;  // children share the parent's panel
    int cand[3][64];
    int ncand[3] = {0, 0, 0};
    const int Ns[3] = {N0, N1, N2};
    for (int a = 0; a < 3; a++) {
        for (int q = 0; q < Ns[a] && ncand[a] < 64; q++) {
            int ci = a == 0 ? q * N1 * N2 : (a == 1 ? q * N2 : q);
            const Node &c = nodes[n.first_child + ci];
            if ((h[a] - c.rng[a][0]) * (h[a] - c.rng[a][1]) <= 0.0) cand[a][ncand[a]++] = q;
        }
    }
    for (int x = 0; x < ncand[0]; x++)
        for (int y = 0; y < ncand[1]; y++)
            for (int z = 0; z < ncand[2]; z++) {
                int ch = n.first_child + (cand[0][x] * N1 + cand[1][y]) * N2 + cand[2][z];
                int r = -777;
                if (inside(ch, rtp, rtp_other, yin)) r = find_branch(ch, rtp, rtp_other, yin);
                if (r != -777) return r;
            }
    return -777;
}

// YinYangTree_FindNode (3_6/ModYinYangTree.f90:449-464): Yin root tested first.
int Tree::find(const double rtp[3], int yin) const {
    double other[3];
    coordconv(rtp, other);
    if (inside(0, rtp, other, yin)) return find_branch(0, rtp, other, yin);
    return find_branch(1, rtp, other, yin);
}

static void init_geom(const Tree &t, const Node &n, BlockGeom &g) {
    const int ni = t.ni, nj = t.nj, nk = t.nk, ng = t.ng;
    g.yin = n.yin;
    std::memcpy(g.rng, n.rng, sizeof(g.rng));
    const int nn[3] = {ni, nj, nk};
    std::vector<double> *x[3] = {&g.xi, &g.xj, &g.xk};
    // ModBlock_InitGrid (3_6/ModBlock.f90:310-317)
    for (int a = 0; a < 3; a++) {
        x[a]->resize(nn[a] + 2 * ng);
        for (int i = 1 - ng; i <= nn[a] + ng; i++)
            (*x[a])[i + ng - 1] = (n.rng[a][0] * (double)((float)(nn[a] - i) + 0.5f) +
                                   n.rng[a][1] * (double)((float)i - 0.5f)) / (double)nn[a];
    }
    g.dxi = (n.rng[0][1] - n.rng[0][0]) / (double)ni;
    g.dxj = (n.rng[1][1] - n.rng[1][0]) / (double)nj;
    g.dxk = (n.rng[2][1] - n.rng[2][0]) / (double)nk;
    // YinYangTree_if_top_bottom (3_6/ModYinYangTree.f90:273-288)
    g.top = g.xi[ni + 2 * ng - 1] > t.r_range[1];
    g.bottom = g.xi[0] < t.r_range[0];
}

void Layout::build_geom() {
    if (tree.nleaf == 0) tree.label();
    const int nb = tree.nleaf;
    geom.assign(nb, BlockGeom());
    for (int b = 0; b < nb; b++) init_geom(tree, tree.nodes[tree.leaf_node[b]], geom[b]);
}

void conv_rows(const Layout &lay, int nthreads, std::vector<double> &rows) {
    const int nb = lay.tree.nleaf, ng = lay.tree.ng;
    const int NJ = lay.tree.nj + 2 * ng, NK = lay.tree.nk + 2 * ng;
    rows.assign((size_t)nb * NJ * NK * 4, 0.0);
    auto work = [&](int b) {
        const BlockGeom &g = lay.geom[b];
        for (int J = 0; J < NJ; J++)
            for (int K = 0; K < NK; K++) {
                // as Tree::find and the Yin-Yang entries of build() form them:
                // the point in the other frame, then the rotation of that
                // (donor-frame) point, whose CoordConv leads back
                const double x[3] = {g.xi[0], g.xj[J], g.xk[K]};
                double o[3], o2[3];
                coordconv(x, o);
                coordconv(o, o2);
                double sp, cp;
                ::sincos(o[2], &sp, &cp);
                double *r = &rows[(((size_t)b * NJ + J) * NK + K) * 4];
                r[0] = o[1];
                r[1] = o[2];
                r[2] = sp * std::sin(o2[2]);
                r[3] = cp / std::sin(o2[1]);
            }
    };
    std::vector<std::thread> th;
    const int nt = std::max(1, std::min(nthreads, nb));
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t]() { for (int b = t; b < nb; b += nt) work(b); });
    for (auto &x : th) x.join();
}

void Layout::build(int nthreads) {
    build_geom();
    const int nb = tree.nleaf;
    const int ni = tree.ni, nj = tree.nj, nk = tree.nk, ng = tree.ng;
    const int NI = ni + 2 * ng, NJ = nj + 2 * ng, NK = nk + 2 * ng;
    targets.assign(nb, {});
    if (keep_gcib) gcib.assign(nb, {});
    std::vector<std::vector<GhostEntry>> per(nb);
    if (nthreads < 1) nthreads = 1;

    auto work = [&](int b) {
        const BlockGeom &g = geom[b];
        std::vector<int32_t> donor((size_t)NI * NJ * NK, -777);
        // ModGC_FindGC_Single (3_6/ModGC.f90:37-50): i outer, k inner
        for (int i = 1 - ng; i <= ni + ng; i++)
            for (int j = 1 - ng; j <= nj + ng; j++)
                for (int k = 1 - ng; k <= nk + ng; k++) {
                    if (i < 1 || j < 1 || k < 1 || i > ni || j > nj || k > nk) {
                        double rtp[3] = {g.xi[i + ng - 1], g.xj[j + ng - 1], g.xk[k + ng - 1]};
                        donor[(size_t)(i + ng - 1) + (size_t)NI * ((j + ng - 1) + (size_t)NJ * (k + ng - 1))] = tree.find(rtp, g.yin);
                    }
                }
        // ModGC_GetGC_Targets_single (3_6/ModGC.f90:53-147): donors in first-appearance order
        auto &tg = targets[b];
        std::vector<int> slot(nb + 1, -1);
        for (int i = 1 - ng; i <= ni + ng; i++)
            for (int j = 1 - ng; j <= nj + ng; j++)
                for (int k = 1 - ng; k <= nk + ng; k++) {
                    if (!(i < 1 || i > ni || j < 1 || j > nj || k < 1 || k > nk)) continue;
                    size_t q = (size_t)(i + ng - 1) + (size_t)NI * ((j + ng - 1) + (size_t)NJ * (k + ng - 1));
                    int d = donor[q];
                    if (d == -777) continue;
                    if (slot[d] < 0) { slot[d] = (int)tg.size(); tg.push_back({d, {}}); }
                    tg[slot[d]].second.push_back((int32_t)q);
                }
        // interpolation base and weights (3_6/ModMath.f90:265-280) in the donor frame
        auto &out = per[b];
        for (auto &t : tg) {
            const int d = t.first - 1;
            const BlockGeom &s = geom[d];
            const bool yy = s.yin != g.yin;
            const double dxi = s.xi[ng + 1] - s.xi[ng], dxj = s.xj[ng + 1] - s.xj[ng], dxk = s.xk[ng + 1] - s.xk[ng];
            const double base = (double)((float)(-ng) + 1.0f);
            const double lim[3] = {(double)((float)ni + 0.0f), (double)((float)nj + 0.0f), (double)((float)nk + 0.0f)};
            for (int32_t q : t.second) {
                int gi = (int)(q % NI) - ng + 1, gj = (int)((q / NI) % NJ) - ng + 1, gk = (int)(q / ((size_t)NI * NJ)) - ng + 1;
                double x[3] = {g.xi[gi + ng - 1], g.xj[gj + ng - 1], g.xk[gk + ng - 1]};
                if (yy) { double o[3]; coordconv(x, o); x[0] = o[0]; x[1] = o[1]; x[2] = o[2]; }
                double p[3] = {base + (x[0] - s.xi[0]) / dxi, base + (x[1] - s.xj[0]) / dxj, base + (x[2] - s.xk[0]) / dxk};
                GhostEntry e;
                e.recv_block = b; e.gi = gi; e.gj = gj; e.gk = gk; e.donor = d;
                int32_t base_idx[3];
                for (int a = 0; a < 3; a++) {
                    double v = p[a] < lim[a] ? p[a] : lim[a];
                    v = v > 1.0 ? v : 1.0;
                    base_idx[a] = (int32_t)std::floor(v);
                    e.w[a] = v - (double)base_idx[a];
                }
                e.i0 = base_idx[0]; e.j0 = base_idx[1]; e.k0 = base_idx[2];
                e.yy = yy;
                e.rot[0] = e.rot[1] = 0.0;
                if (yy) {
                    // ModYinYang_VecConv_1D (3_6/ModYinYang.f90:48-61) coefficients
                    double o[3];
                    coordconv(x, o);
                    double sp, cp;
                    ::sincos(x[2], &sp, &cp);
                    e.rot[0] = sp * std::sin(o[2]);
                    e.rot[1] = cp / std::sin(o[1]);
                }
                out.push_back(e);
            }
        }
        if (keep_gcib) {
            for (auto &v : donor) if (v != -777 && v < 0) v = -777;
            gcib[b] = std::move(donor);
        }
    };

    std::vector<std::thread> th;
    int nt = std::min(nthreads, nb);
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t]() { for (int b = t; b < nb; b += nt) work(b); });
    for (auto &x : th) x.join();

    size_t total = 0;
    for (auto &v : per) total += v.size();
    ghosts.clear();
    ghosts.reserve(total);
    for (auto &v : per) { ghosts.insert(ghosts.end(), v.begin(), v.end()); std::vector<GhostEntry>().swap(v); }
}

void polytrope(const PolytropeParams &p, double r, double out[8]) {
    const double n = 1.5, G = 6.674e-8, M = 1.989e33, kB = 1.380649e-16, mp = 1.67262192369e-24;
    const double zeta_lo = std::pow(p.contrast, 1.0 / n);
    const double beta = (zeta_lo - 1.0) / (p.r_hi / p.r_lo - 1.0);
    const double zeta = 1.0 + beta * (p.r_hi / r - 1.0);
    const double pr = G * M / ((n + 1.0) * beta * p.r_hi);  // p_b / rho_b
    const double rho = p.rho_top * std::pow(zeta, n);
    const double pres = p.rho_top * pr * std::pow(zeta, n + 1.0);
    out[0] = (n + 1.0) * pr * beta * p.r_hi / (r * r);  // = G M / r^2, hydrostatic with the profile
    out[1] = rho;
    out[2] = pres;
    out[3] = p.mu * mp / kB * pr * zeta;
    out[4] = 5.0 / 3.0;
    out[5] = 5.0 / 3.0;
    out[6] = 0.0;
    out[7] = 1.0;
}

}  // namespace falcon

// ---------------------------------------------------------------------------
// Heat-conduction levels.  Face areas, volumes and centre distances follow
// ModMultigrid_InitGrid / ModBlock_InitGrid (3_5/ModMultigrid.f90:150-233,
// 3_6/ModBlock.f90:367-415); the operator follows the finite-volume assembly of
// ModPFSS_setup (3_5/ModPFSS.f90:68-84) with kappa(r) at the faces.
namespace falcon {

namespace {
struct LevelGeom {
    double rng[3][2];
    int yin;
    std::vector<double> xc[3];   // padded centres, [0] <-> index 0 (one ghost layer)
    std::vector<double> xf[3];   // faces 1..m+1 -> [0..m]
    double dx[3];
};

LevelGeom level_geom(const Node &n, int m) {
    LevelGeom g;
    std::memcpy(g.rng, n.rng, sizeof(g.rng));
    g.yin = n.yin;
    for (int a = 0; a < 3; a++) {
        g.xc[a].resize(m + 2);
        g.xf[a].resize(m + 1);
        for (int i = 0; i <= m + 1; i++)
            g.xc[a][i] = (n.rng[a][0] * (double)((float)(m - i) + 0.5f) + n.rng[a][1] * (double)((float)i - 0.5f)) / (double)m;
        for (int i = 1; i <= m + 1; i++)
            g.xf[a][i - 1] = (n.rng[a][0] * (double)(m + 1 - i) + n.rng[a][1] * (double)(i - 1)) / (double)m;
        g.dx[a] = (n.rng[a][1] - n.rng[a][0]) / (double)m;
    }
    return g;
}
}  // namespace

void build_pfss_tables(const Layout &lay, PfssTables &pt) {
    const Tree &t = lay.tree;
    const int m = t.ni, nb = t.nleaf;
    pt.m = m;
    pt.bottom.clear();
    pt.area.clear();
    pt.dr.assign(nb, 0.0);
    pt.geo.assign((size_t)nb * (2 * m + 3), 0.0);
    for (int b = 0; b < nb; b++) {
        const LevelGeom g = level_geom(t.nodes[t.leaf_node[b]], m);
        pt.dr[b] = g.dx[0];
        double *G = &pt.geo[(size_t)b * (2 * m + 3)];
        for (int i = 0; i < m; i++) G[i] = 1.0 / g.xc[0][i + 1];
        for (int j = 0; j < m; j++) G[m + j] = 1.0 / std::sin(g.xc[1][j + 1]);
        for (int a = 0; a < 3; a++) G[2 * m + a] = (double)(1.0f / 2.0f) / g.dx[a];
        if (!lay.geom[b].bottom) continue;
        pt.bottom.push_back(b);
        const double rf = g.xf[0][0], dph = g.dx[2];
        for (int k = 0; k < m; k++)
            for (int j = 0; j < m; j++)
                pt.area.push_back(rf * rf * (std::cos(g.xf[1][j]) - std::cos(g.xf[1][j + 1])) * dph);
    }
}

void dipole_magnetogram(const Layout &lay, const PfssTables &pt, int n, const double *xyz, const double *q,
                        const double *l, std::vector<double> &br) {
    const Tree &t = lay.tree;
    const int m = pt.m;
    br.assign(pt.bottom.size() * (size_t)m * m, 0.0);
    for (size_t bb = 0; bb < pt.bottom.size(); bb++) {
        const int b = pt.bottom[bb];
        const LevelGeom g = level_geom(t.nodes[t.leaf_node[b]], m);
        const double rF = g.xf[0][0];
        for (int j = 0; j < m; j++)
            for (int k = 0; k < m; k++) {
                // rtp_face_center = [xi_F(1), xj_F(j), xk_I(k)]: theta at the face
                // xj_F(j) and phi at the centre, as the reference writes it; Yin
                // points in the other panel's frame
                double rtp[3] = {rF, g.xf[1][j], g.xc[2][k + 1]};
                if (g.yin) {
                    double o[3];
                    coordconv(rtp, o);
                    rtp[0] = o[0]; rtp[1] = o[1]; rtp[2] = o[2];
                }
                double sth, cth, sph, cph;
                ::sincos(rtp[1], &sth, &cth);
                ::sincos(rtp[2], &sph, &cph);
                const double f[3] = {rtp[0] * (cph * sth), rtp[0] * (sph * sth), rtp[0] * cth};
                double o = 0.0;
                for (int d = 0; d < n; d++) {
                    double dp[3], dn[3];
                    for (int a = 0; a < 3; a++) {
                        dp[a] = f[a] - (xyz[3 * d + a] + l[3 * d + a] / 2.0);
                        dn[a] = f[a] - (xyz[3 * d + a] - l[3 * d + a] / 2.0);
                    }
                    const double rp = std::sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
                    const double rn = std::sqrt(dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2]);
                    double sp = 0.0, sn = 0.0;
                    for (int a = 0; a < 3; a++) { sp = sp + dp[a] * f[a] / rF; sn = sn + dn[a] * f[a] / rF; }
                    // the radial component of the two poles' Coulomb fields
                    o = (o + q[d] * sp / (rp * rp * rp)) - q[d] * sn / (rn * rn * rn);
                }
                br[bb * (size_t)m * m + (size_t)k * m + j] = o;
            }
    }
}

void build_heat_levels(const Layout &lay, const HeatParams &hp, int nthreads, std::vector<HeatLevel> &levels) {
    const Tree &t = lay.tree;
    if (t.ni != t.nj || t.ni != t.nk) throw std::runtime_error("heat levels need cubic blocks");
    int n = t.ni;
    if (n & (n - 1)) throw std::runtime_error("heat levels need a power-of-two block edge");
    const int nb = t.nleaf;
    levels.clear();
    for (int m = n; m >= 1; m /= 2) {
        HeatLevel L;
        L.m = m;
        const int M2 = m * m, M3 = m * m * m;
        L.V.assign((size_t)nb * M2, 0.0);
        L.offsum.assign((size_t)nb * M2, 0.0);
        L.Ai.assign((size_t)nb * (m + 1) * m, 0.0);
        L.Aj.assign((size_t)nb * m * (m + 1), 0.0);
        L.Ak.assign((size_t)nb * M2, 0.0);
        L.g_src.assign((size_t)nb * 6 * M2, -1);
        L.g_wid.assign((size_t)nb * 6 * M2, -1);
        std::vector<std::vector<double>> gw(nb);
        std::vector<std::vector<int32_t>> gwslot(nb);     // ghost ordinal of each weighted entry
        if (hp.kappa_power != -3 && hp.kappa_power != 0) throw std::runtime_error("kappa_power must be -3 or 0");
        auto kap = [&](double r) {
            if (hp.kappa_power == 0) return hp.kappa0;
            double x = r / hp.r_ref;
            return hp.kappa0 / (x * x * x);
        };
        auto work = [&](int b) {
            const Node &nd = t.nodes[t.leaf_node[b]];
            LevelGeom g = level_geom(nd, m);
            const bool top = lay.geom[b].top, bottom = lay.geom[b].bottom;
            const double *rf = g.xf[0].data(), *tf = g.xf[1].data();
            const double *rc = g.xc[0].data(), *tc = g.xc[1].data();
            const double dph = g.dx[2];
            double *V = &L.V[(size_t)b * M2], *off = &L.offsum[(size_t)b * M2];
            double *Ai = &L.Ai[(size_t)b * (m + 1) * m], *Aj = &L.Aj[(size_t)b * m * (m + 1)], *Ak = &L.Ak[(size_t)b * M2];
            for (int i = 0; i <= m; i++)
                for (int j = 0; j < m; j++) {
                    double S = rf[i] * rf[i] * (std::cos(tf[j]) - std::cos(tf[j + 1])) * dph;
                    double c = kap(rf[i]) * S / g.dx[0];
                    // insulated radial boundary (the PFSS bottom too: its flux is the
                    // right-hand side); the top is insulated or Dirichlet
                    if ((i == 0 && bottom) || (i == m && top && !hp.top_dirichlet)) c = 0.0;
                    Ai[i * m + j] = c;
                }
            for (int i = 0; i < m; i++)
                for (int j = 0; j <= m; j++) {
                    double S = 0.5 * (rf[i + 1] * rf[i + 1] - rf[i] * rf[i]) * dph * std::sin(tf[j]);
                    double D = rc[i + 1] * g.dx[1];
                    Aj[i * (m + 1) + j] = kap(rc[i + 1]) * S / D;
                }
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) {
                    double S = 0.5 * (rf[i + 1] * rf[i + 1] - rf[i] * rf[i]) * (tf[j + 1] - tf[j]);
                    double D = rc[i + 1] * g.dx[2] * std::sin(tc[j + 1]);
                    Ak[i * m + j] = kap(rc[i + 1]) * S / D;
                }
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) {
                    V[i * m + j] = 1.0 / 3.0 * (rf[i + 1] * rf[i + 1] * rf[i + 1] - rf[i] * rf[i] * rf[i]) *
                                   (std::cos(tf[j]) - std::cos(tf[j + 1])) * dph;
                    off[i * m + j] = Ai[i * m + j] + Ai[(i + 1) * m + j] + Aj[i * (m + 1) + j] + Aj[i * (m + 1) + j + 1] +
                                     Ak[i * m + j] + Ak[i * m + j];
                    // Dirichlet top face: Aii = Aii - Ai(top) with Ai(top) = -S/D
                    // (3_5/ModPFSS.f90:88-91), i.e. the face once more
                    if (hp.top_dirichlet && top && i == m - 1) off[i * m + j] = off[i * m + j] + Ai[m * m + j];
                }
            // face ghosts in face order i-, i+, j-, j+, k-, k+; radial domain boundary
            // faces (no donor) keep src = -1 and a zero coefficient
            for (int f = 0; f < 6; f++)
                for (int t1 = 0; t1 < m; t1++)
                    for (int t2 = 0; t2 < m; t2++) {
                        const int ax = f / 2, side = f % 2;
                        int P[3];
                        const int tt[2] = {t1 + 1, t2 + 1};
                        for (int a = 0, u = 0; a < 3; a++) P[a] = a == ax ? (side ? m + 1 : 0) : tt[u++];
                        double x[3] = {g.xc[0][P[0]], g.xc[1][P[1]], g.xc[2][P[2]]};
                        int d = t.find(x, g.yin);
                        if (d == -777) continue;
                        d -= 1;
                        const LevelGeom dg = level_geom(t.nodes[t.leaf_node[d]], m);
                        if (dg.yin != g.yin) { double o[3]; coordconv(x, o); x[0] = o[0]; x[1] = o[1]; x[2] = o[2]; }
                        // linear inter/extrapolation inside the donor block: the point's padded
                        // index p lies in [0.5, m+0.5]; the stencil base is kept in [1, m-1] so
                        // only interior donor cells are read and fields linear in (r, theta, phi)
                        // are reproduced exactly across blocks of different spacing.  m = 1
                        // copies the donor's single cell.  Weights within 1e-12 of 0 or 1 are
                        // snapped (the rounding residue of coincident cell centres).
                        int base[3];
                        double w[3];
                        bool copy = true;
                        for (int a = 0; a < 3; a++) {
                            if (m == 1) { base[a] = 1; w[a] = 0.0; continue; }
                            const double dxa = dg.xc[a][2] - dg.xc[a][1];
                            double p = (x[a] - dg.xc[a][0]) / dxa;    // padded index of the point
                            p = p < (double)m + 0.5 ? p : (double)m + 0.5;
                            p = p > 0.5 ? p : 0.5;
                            int bs = (int)std::floor(p);
                            bs = bs < 1 ? 1 : (bs > m - 1 ? m - 1 : bs);
                            base[a] = bs;
                            double wa = p - (double)bs;
                            if (std::fabs(wa) < 1e-12) wa = 0.0;
                            else if (std::fabs(wa - 1.0) < 1e-12) wa = 1.0;
                            w[a] = wa;
                            copy = copy && (wa == 0.0 || wa == 1.0);
                        }
                        const size_t gi = ((size_t)b * 6 + f) * M2 + (size_t)t1 * m + t2;
                        long long src = (long long)d * M3 + ((long long)(base[0] - 1) * m + (base[1] - 1)) * m + (base[2] - 1);
                        if (copy) {
                            src += (long long)w[0] * M2 + (long long)w[1] * m + (long long)w[2];
                        } else {
                            gwslot[b].push_back((int32_t)gi);
                            gw[b].push_back(w[0]); gw[b].push_back(w[1]); gw[b].push_back(w[2]);
                        }
                        L.g_src[gi] = (int32_t)src;
                    }
        };
        std::vector<std::thread> th;
        int nt = std::max(1, std::min(nthreads, nb));
        for (int w = 0; w < nt; w++) th.emplace_back([&, w]() { for (int b = w; b < nb; b += nt) work(b); });
        for (auto &x : th) x.join();
        for (int b = 0; b < nb; b++) {
            for (size_t e = 0; e < gwslot[b].size(); e++) L.g_wid[gwslot[b][e]] = (int32_t)(L.g_w.size() / 3 + e);
            L.g_w.insert(L.g_w.end(), gw[b].begin(), gw[b].end());
        }
        for (size_t gi = 0; gi < L.g_src.size(); gi++) {
            if (L.g_src[gi] < 0) L.n_none++;
            else if (L.g_wid[gi] < 0) L.n_copy++;
            else L.n_weighted++;
        }
        levels.push_back(std::move(L));
    }
}

}  // namespace falcon
