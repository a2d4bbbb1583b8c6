// Shared declarations between the kernel file and the C-ABI file.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace falcon {

struct Layout;
// Layout::build's ghost tables (donors, targets, entries, and GC_iBlocks when
// the layout keeps them) built on the device (tables.hip); equal to the host
// build bit for bit
void build_ghosts_device(Layout &lay, int nthreads, hipStream_t s);

// Per-block constants, FP64 (3_6/ModBlock.f90:315-317 and the derivative
// coefficient (1./12./dx) of 3_6/ModDerivative.f90:16,22,28).
struct BlockMeta {
    double cr, ct, cp;      // (double)(1.0f/12.0f)/dx
    double dxi, dxj, dxk;
    int flags;              // bit 0: bottom block (inner radial boundary), bit 1: top block
};

// Geometry and background tables on the device, one row per block.
// Radial rows have NI entries, angular rows NJ entries (padded indices).
struct DevTables {
    const BlockMeta *meta;
    // radial (NI per block)
    const double *xi, *rinv, *r2inv, *rho0, *gp, *gq, *g_over_rho0, *xir2, *heat, *cs, *sqrt_rho0;
    // theta (NJ per block)
    const double *sint, *sinv, *cot;
};

struct Dims {
    int ni, nj, nk, ng;
    int NI, NJ, NK;
    int nblocks;
    int tiles_j, tiles_k;   // 16x16 (theta, phi) tiles per block (k_dt_tile and the 16-edge marches)
    long long P;            // cells per variable per block (NI*NJ*NK)
    // cell layout within one block variable (cell_off): main part strides
    int SJ;                 // nk
    long long SI;           // NJ*nk
    long long M;            // NI*NJ*nk, the first cell of the side part
};

// Offset of padded cell (I, J, K) within one block variable of a state buffer
// (buffers are [block][variable][P cells]).  The P cells are split in two
// parts so that every row a march reads or writes is whole and aligned:
// - main: the interior phi range of every padded (I, J) row, rows of nk
//   values: (I*NJ + J)*nk + (K - ng).  A 16^3 block's rows are 64 (FP32) or
//   128 (FP64) bytes, and a wavefront's four rows are one contiguous span;
// - side: the 2*ng phi ghosts of every row: M + (I*NJ + J)*2*ng + kg, with
//   kg = K below the interior and K - nk above it.
// Rows padded in place (20 values with the interior at 2..17) made every
// store of a wavefront a set of partial 128-byte lines: such stores measured
// 1.56 TB/s on MI355X against 5.9 TB/s for whole aligned rows.
__host__ __device__ inline long long cell_off(const Dims &d, int I, int J, int K) {
    if (K >= d.ng && K < d.ng + d.nk) return ((long long)I * d.NJ + J) * d.nk + (K - d.ng);
    return d.M + ((long long)I * d.NJ + J) * (2 * d.ng) + (K < d.ng ? K : K - d.nk);
}
// inverse of cell_off
__host__ __device__ inline void cell_ijk(const Dims &d, long long c, int &I, int &J, int &K) {
    if (c < d.M) {
        K = (int)(c % d.nk) + d.ng;
        c /= d.nk;
    } else {
        c -= d.M;
        const int kg = (int)(c % (2 * d.ng));
        K = kg < d.ng ? kg : kg + d.nk;
        c /= 2 * d.ng;
    }
    J = (int)(c % d.NJ);
    I = (int)(c / d.NJ);
}
// Dims of a layout of padded blocks (ni, nj, nk, ng)
inline Dims make_dims(int ni, int nj, int nk, int ng, int nblocks) {
    Dims d{};
    d.ni = ni; d.nj = nj; d.nk = nk; d.ng = ng;
    d.NI = ni + 2 * ng; d.NJ = nj + 2 * ng; d.NK = nk + 2 * ng;
    d.nblocks = nblocks;
    d.tiles_j = (nj + 15) / 16;
    d.tiles_k = (nk + 15) / 16;
    d.P = (long long)d.NI * d.NJ * d.NK;
    d.SJ = nk;
    d.SI = (long long)d.NJ * nk;
    d.M = (long long)d.NI * d.NJ * nk;
    return d;
}

// Ghost-fill tables (device), one row per ghost entry: same-panel entries in
// [0, yy_begin), Yin-Yang entries in [yy_begin, n_total), each set in the order
// of the receiving blocks' target lists.
struct DevGhosts {
    const int32_t *dst;     // block*8*P + cell of the ghost (variable 0)
    const int32_t *src;     // donor*8*P + cell of the interpolation base (variable 0)
    const double *w;        // 3 weights per entry
    const double *rot;      // 2 rotation coefficients per entry (read for Yin-Yang entries only)
    int first, n_total, yy_begin;  // weighted entries [first, n_total): same-panel, then Yin-Yang from yy_begin
    const int2 *copy;       // (dst, src) of the entries filled by a plain copy, sorted by dst
    int n_copy;
    const int2 *lerp;       // (dst, src) of the entries interpolated along r only, sorted by dst
    const double *lerp_w;   // their r weight
    int n_lerp;
    // copies whose ghosts the consuming kernels read in place (see pull below),
    // sorted by dst
    const int2 *copy_rest;
    int n_copy_rest;
    // pull[R*27 + (di+1)*9 + (dj+1)*3 + (dk+1)]: the same-level block D whose
    // interior holds all ghosts of block R in direction (di,dj,dk) as plain
    // copies, ghost index g <- D's cell g - (di*ni, dj*nj, dk*nk); -1 where any
    // ghost of that region is filled otherwise.  null: no pull.
    const int32_t *pull;
};

struct DevBoundary {
    const int32_t *list;    // block*2 + side (0 bottom, 1 top)
    int n;
};

// launchers (kernels.hip)
// pull (null: none): the pull table of DevGhosts; ghosts of `in` in the
// directions it lists are read from the neighbour's interior instead.
// tile: (theta, phi) tile edge of the march, 8 or 16; wgs: 0 launches one
// workgroup per tile, n > 0 launches n workgroups that loop over the tiles (a
// persistent launch).  Results depend on neither; the dt partial minima are
// laid out per tile of the diffusion's edge.
// glm (FP64 only; null: off): grad(psi) per interior cell, [block][3][ni*nj*nk],
// subtracted from R(B) before the update (GLM divergence cleaning)
template <typename T, typename TIN, typename TOUT, typename TB>
void launch_stage(const Dims &d, const DevTables &t, const TIN *in, const TB *U, TOUT *out, float *shadow,
                  const double *dtp, double div, hipStream_t s, const int32_t *pull, int tile, int wgs,
                  const double *glm = nullptr);
// GLM divergence cleaning (FP64 path).  psi buffers are [block][P] in the
// cell_off layout.  Speed (once per step): Ch = c_h and c_h / h_LLL of the
// step's state U, [block][2][interior cells].  Stage: psi_out = psi_n +
// dt R(psi)/div from the stage input (state `in`, psi_in) and Ch, and
// G = grad(psi_in);
// diffusion: psi_new = psi + 0.5 dt R_diff(psi) with the speeds of U;
// ghost fill: the FP64 state's and psi's ghosts in one pass over the FP64
// ghost tables (psi a scalar: no Yin-Yang rotation).
void launch_glm_speed(const Dims &d, const DevTables &t, const double *U, double *Ch, hipStream_t s);
void launch_glm_stage(const Dims &d, const DevTables &t, const double *in, const double *Ch, const double *psi_in,
                      const double *psi_n, double *psi_out, double *G, const double *dtp, double div, hipStream_t s);
void launch_glm_diffusion(const Dims &d, const DevTables &t, const double *U, const double *psi, double *psi_new,
                          const double *dtp, hipStream_t s);
void launch_exchange_glm(const Dims &d, const DevGhosts &g, double *state, double *psi, hipStream_t s);
// Uold: U^n for the FP32 path, whose diffusion updates U (= out) in place;
// null for the FP64 path (U != out)
template <typename T, typename TIN>
void launch_diffusion(const Dims &d, const DevTables &t, const TIN *in, const double *U, double *out, float *shadow,
                      double *partial, const double *dtp, hipStream_t s, const int32_t *pull, int tile, int wgs,
                      const double *Uold = nullptr);
// workgroups of the stage (kernel 0) or diffusion (1) march of a precision that
// the device holds resident at once: the persistent grid
int march_resident_wgs(int kernel, int precision, int tile);
// every ghost entry of g
template <typename TS>
void launch_exchange(const Dims &d, const DevGhosts &g, TS *state, hipStream_t s);
// the entries that kernels reading with pull = g.pull still need
template <typename TS>
void launch_exchange_rest(const Dims &d, const DevGhosts &g, TS *state, hipStream_t s);
template <typename TS>
void launch_boundary(const Dims &d, const DevBoundary &b, TS *state, hipStream_t s, const int32_t *pull = nullptr);
template <typename TS>
void launch_copy_radial(const Dims &d, const DevBoundary &b, const TS *src, TS *dst, hipStream_t s);
void launch_dt_partial(const Dims &d, const DevTables &t, const double *U, double *partial, hipStream_t s, int tile);
void launch_dt_final(const Dims &d, const double *partial, double *dtp, double cfl, hipStream_t s, int tile);
void launch_to_f32(const double *src, float *dst, long long n, hipStream_t s);
void launch_set_scalar(double *p, double v, hipStream_t s);
// NaN scan of a padded state (see k_nan_scan): w[0] counts, w[1] first position;
// w starts as {0, ~0}
template <typename TS>
void launch_nan_scan(const Dims &d, const TS *st, unsigned long long *w, hipStream_t s);
// the same scan of psi ([block][P]) as variable 8 (GLM divergence cleaning)
void launch_nan_scan_psi(const Dims &d, const double *psi, unsigned long long *w, hipStream_t s);
// interior difference of two states (see k_state_diff): pass 0 then pass 1 on
// the same 32 words w, whose first 8 and last 8 start at 0 and at ~0
void launch_state_diff(const Dims &d, const double *a, const double *b, double tol, unsigned long long *w, int pass,
                       hipStream_t s);
// one Fourier mode of the synthetic vector potential: B contribution
// a cos(k.x + ph), a = k x (random amplitude)
struct SynthMode {
    double k[3], a[3], ph;
};
// seeded synthetic initial state on the interior (k_synthetic); rng: 6 per
// block, rows: c_p, sqrt(rho0), c_s (NI each) per block; wgmax: one double
// per 256 interior cells
void launch_synthetic(const Dims &d, const double *rng, const int32_t *yin, const double *rows, const SynthMode *modes,
                      int nm, unsigned long long seed, double s_amp, double b_ratio, double *U, double *wgmax,
                      hipStream_t s);
// corner values of the AllCells output for blocks [b0, b0 + nb) (k_allcells)
void launch_allcells(const Dims &d, const double *U, const double *geo, const int32_t *yin, int b0, int nb,
                     double rsun, double *out, hipStream_t s);
// single-operator entry points used by parity tests: R only
template <typename T, typename TIN>
void launch_rhs_only(const Dims &d, const DevTables &t, const TIN *in, double *R, hipStream_t s, int tile, int wgs);
template <typename T>
void launch_diffusion_only(const Dims &d, const DevTables &t, const double *U, double *R, hipStream_t s, int tile,
                           int wgs);

}  // namespace falcon
