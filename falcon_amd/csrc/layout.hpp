// Host-side layout: Yin-Yang block tree, per-block geometry, background
// profiles and the ghost-fill tables that drive the device exchange kernel.
//
// Tree semantics follow 3_6/ModYinYangTree.f90 (roots :91-110, split rule
// :239-255, depth-first leaf labelling Yin then Yang :294-310/:402-420, point
// location :424-505).  Ghost tables follow 3_6/ModGC.f90:25-147 (donor per
// ghost, targets in first-appearance order) and the interpolation base/weights
// of 3_6/ModMath.f90:250-293 evaluated once on the host.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace falcon {

constexpr int kNVar = 8;  // rho1, vr, vt, vp, br, bt, bp, s1 (3_6/ModBlock.f90:153-163 without psi)

struct Node {
    int yin = 1;
    int leaf = 1;
    int leaf_id = 0;       // 1-based once labelled
    int first_child = -1;  // index into Tree::nodes
    int N[3] = {1, 1, 1};  // children per axis (children ordered r-major, phi-minor)
    double rng[3][2] = {};
};

struct Tree {
    int ni = 0, nj = 0, nk = 0, ng = 2;
    double r_range[2] = {0, 0};
    std::vector<Node> nodes;   // [0] Yin root, [1] Yang root
    int nleaf = 0, nleaf_yin = 0;
    std::vector<int> leaf_node;  // leaf id - 1 -> node index

    void init(int ni_, int nj_, int nk_, int ng_, double r_lo_cgs, double r_hi_cgs);
    void divide(int node, const int N[3]);
    void divide_all(int nr, int nt, int np);
    void amr(const int flags[3], int global, const double rng_rsun[6]);
    void refine(const std::vector<int> &ordinals);
    void label();
    // Leaf id (1-based) of the node containing rtp given in panel `yin`, or -777.
    int find(const double rtp[3], int yin) const;

  private:
    int find_branch(int node, const double rtp[3], const double rtp_other[3], int yin) const;
    bool inside(int node, const double rtp[3], const double rtp_other[3], int yin) const;
};

struct BlockGeom {
    int yin = 1, top = 0, bottom = 0;
    double rng[3][2] = {};
    double dxi = 0, dxj = 0, dxk = 0;
    std::vector<double> xi, xj, xk;  // padded centres, [0] <-> index -ng+1
};

// One ghost cell of one receiving block, as it is filled from its donor.
struct GhostEntry {
    int32_t recv_block;   // 0-based
    int32_t gi, gj, gk;   // ghost indices in the receiving block (Fortran, -ng+1..n+ng)
    int32_t donor;        // 0-based
    int32_t i0, j0, k0;   // donor base cell (Fortran indices, clamped to [1,n])
    double w[3];          // trilinear weights
    int32_t yy;           // 1 if donor and receiver are in different panels
    double rot[2];        // a = sin(p)sin(p'), b = cos(p)/sin(t') (donor-frame point)
};

struct Layout {
    Tree tree;
    std::vector<BlockGeom> geom;
    // per block: ordered list of (donor leaf id, [ghost ordinal...]) = GC_targets
    std::vector<std::vector<std::pair<int32_t, std::vector<int32_t>>>> targets;
    std::vector<GhostEntry> ghosts;        // grouped by receiving block, then target order
    std::vector<std::vector<int32_t>> gcib; // optional: GC_iBlocks_III per block
    bool keep_gcib = false;

    int nblocks() const { return tree.nleaf; }
    // per-block geometry of every leaf (labels the tree if needed)
    void build_geom();
    // geometry and ghost tables on the host
    void build(int nthreads);
};

// Transcendental rows of the device table builder, per block [J][K] over the
// padded (theta, phi) centres, 4 doubles: theta' and phi' of CoordConv
// (3_6/ModYinYang.f90:29-46, the point in the other panel's frame) and the
// VecConv coefficients (:48-61) of that donor-frame point,
// sin(phi') sin(phi''), cos(phi') / sin(theta'') with (theta'', phi'') its
// CoordConv.  The device evaluates the search and the weights from them with
// +, -, *, / and floor only, so its tables equal the host's bit for bit.
void conv_rows(const Layout &lay, int nthreads, std::vector<double> &rows);

// ---------------------------------------------------------------- heat levels
// One multigrid level of the heat operator: every block holds m^3 cells.  The
// 7-point operator reads one face ghost per face-adjacent cell; ghost values
// are trilinear inter/extrapolations inside the donor block (weights within
// 1e-12 of 0 or 1 snapped, so same-level neighbours are plain copies),
// evaluated by the consumer from the tables below, never stored.
// Every coefficient depends on (i, j) only: the phi spacing of a block is
// uniform (dphi = (phi_hi - phi_lo) / m), so tables are [block][i][j].
struct HeatLevel {
    int m = 0;                           // cells per block edge
    std::vector<double> V;               // [block][m][m] cell volumes
    std::vector<double> Ai;              // [block][m+1][m] radial faces (0 on insulated boundaries)
    std::vector<double> Aj;              // [block][m][m+1] theta faces
    std::vector<double> Ak;              // [block][m][m] phi faces (both faces of a cell)
    std::vector<double> offsum;          // [block][m][m] sum of the cell's six face coefficients
    // face ghosts: 6 m^2 per block in face order i-, i+, j-, j+, k-, k+, each
    // face [t1][t2] over its tangential axes in (i, j, k) order.  g_src: donor
    // interior base cell (block * m^3 + (i m + j) m + k), -1 without a donor
    // (insulated radial face); g_wid: -1 for a copy, else the index of the 3
    // weights in g_w
    std::vector<int32_t> g_src, g_wid;
    std::vector<double> g_w;
    long long n_copy = 0, n_weighted = 0, n_none = 0;
};

struct HeatParams {
    double kappa0 = 1.0;     // kappa(r) = kappa0 * (r / r_ref)^kappa_power
    double r_ref = 0.0;      // cm
    int kappa_power = -3;    // -3 (heat conduction) or 0 (uniform: the PFSS Laplacian)
    // top radial boundary: insulated (false) or Dirichlet value 0 on the face
    // (true, the PFSS source surface of 3_5/ModPFSS.f90:88-91: the face
    // coefficient enters the diagonal twice, as a mirror ghost -x would, and
    // couples to nothing)
    bool top_dirichlet = false;
};

// Builds levels m = n, n/2, ..., 1 for cubic blocks (ni = nj = nk = n, power of 2).
void build_heat_levels(const Layout &lay, const HeatParams &hp, int nthreads, std::vector<HeatLevel> &levels);

// Host tables of the PFSS potential field on the finest heat level (edge m):
// bottom: the blocks at the domain's inner radius, in block order;
// area: for each of them the areas r_F^2 (cos t_j - cos t_j+1) dphi of its
// bottom faces (j fastest, 3_5/ModMultigrid.f90 Si_FLL(1,:,:));
// dr: every block's radial spacing;
// geo: per block 1/r and 1/sin(theta) of its cell centres (m each), then the
// second-order derivative factors (1./2./dx) of r, theta, phi
// (3_6/ModDerivative.f90:46-56, 3_6/ModSpherical.f90:161-199).
struct PfssTables {
    int m = 0;
    std::vector<int> bottom;
    std::vector<double> area, dr, geo;
};
void build_pfss_tables(const Layout &lay, PfssTables &t);
// ModMagnetogram_dipole_magnetogram_ALL (3_5/ModMagnetogram.f90:13-80): Br of
// n dipoles (centre xyz, pole charge q, pole separation l, cm) at the bottom
// faces of the bottom blocks, (j, k) with j fastest, in FP64 (the reference
// mixes default-real and real(8) operands); br holds nbottom * m * m values.
void dipole_magnetogram(const Layout &lay, const PfssTables &t, int n, const double *xyz, const double *q,
                        const double *l, std::vector<double> &br);

// Yin<->Yang point map (3_6/ModYinYang.f90:8-17); same glibc calls as the reference.
void coordconv(const double in[3], double out[3]);
// sin(t) (a plain libm call, as the reference's operators take it) and
// cos(t)/sin(t) with both from one sincos call (the reference's cotant,
// 3_6/ModSpherical.f90:296, which its build evaluates that way)
void sin_and_cot(double t, double &s, double &cot);

// n=1.5 spherical polytrope background at radius r (cm): writes g, rho0, p0,
// T0, gamma1, gamma3, heat, xi.  rho(r_lo)/rho(r_hi) = contrast.
struct PolytropeParams {
    double r_lo, r_hi, contrast = 20.0, rho_top = 1.0e-2, mu = 0.6;
};
void polytrope(const PolytropeParams &p, double r, double out[8]);

}  // namespace falcon
