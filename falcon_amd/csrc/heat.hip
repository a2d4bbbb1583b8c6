// Implicit heat conduction: multigrid-preconditioned flexible CG on the GPU.
//
// The reference has no heat operator (3_6/ModHeatConduction.f90:15-16 is an
// empty module) and its MG-CG (3_5/ModLinearSolver.f90:313-427) is FP32 dead
// code; this is the method the north-star asks for, defined here:
//   (V/dt) T + sum_f kappa_f S_f/D_f (T - T_nb) = (V/dt) T^n
// on every block's m^3 cells, assembled as ModPFSS_setup assembles its
// Laplacian (3_5/ModPFSS.f90:68-84), kappa(r) = kappa0 (r/r_ref)^-3, insulated
// radial boundaries.  Multigrid levels per block m = 16, 8, 4, 2, 1 with face
// ghosts at every level (the HC exchange of 3_5/ModCommunication.f90:1497-1639,
// here table-driven), summed restriction of the residual, linear
// prolongation from the coarse cell and its face neighbours, damped Jacobi
// (ModMultigrid_Jacobi_Mz_r, 3_5/ModMultigrid.f90:304-325, with damping), and
// on the 1-cell level a Jacobi-preconditioned CG over every block of a system
// (ModLinearSolver_JCB_CG, 3_5/ModLinearSolver.f90:86-180, as a method).
// Krylov: flexible CG (FCG(1)), everything FP64; dot products are
// deterministic two-pass reductions (wavefront shuffle, LDS, second pass).
//
// Layout: every vector holds interior cells only, [system][block][i][j][k].
// Ghost values are never stored: a stencil kernel evaluates the face ghost of
// a boundary cell from the donor table (copy or trilinear) when it needs it,
// so a Jacobi sweep is one launch and reads each vector once.  Coefficients
// are [block][i][j] tables (independent of phi).  Fusions: the first sweep
// from zero is omega r/diag; the residual is restricted in the kernel that
// forms it; dot products are accumulated by the kernels producing their
// operands.
//
// All kernels carry a system index in blockIdx.y: nsys independent systems
// (same grid, own dt and right-hand side) are solved in the same launches, and
// a system that has converged (active[s] = 0) exits every kernel at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/falcon_hip.h"
#include "device.hpp"
#include "heat.hpp"
#include "layout.hpp"

namespace falcon {

#define HCHECK(x)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------ kernels

struct HLev {
    int m, lm, nb;          // cells per edge, log2(m), blocks
    long long M3;           // cells per block
    const double *V, *Ai, *Aj, *Ak, *off;   // [block][i][j] coefficient tables
    const int2 *gsw;        // per face ghost: (donor base cell or -1, weights index or -1 = copy)
    const double *gw;       // 3 weights per weighted ghost
};

struct Cell {
    int b, i, j, k;
};

__device__ __forceinline__ Cell cell_of(const HLev &L, long long c) {
    const int lm = L.lm;
    Cell r;
    r.b = (int)(c >> (3 * lm));
    const int q = (int)(c & (L.M3 - 1));
    r.i = q >> (2 * lm);
    r.j = (q >> lm) & (L.m - 1);
    r.k = q & (L.m - 1);
    return r;
}

// the field a stencil reads
struct Field {
    const double *x;
    __device__ __forceinline__ double operator()(long long c) const { return x[c]; }
};

__device__ __forceinline__ double diag_at(const HLev &L, const Cell &q, double inv_dt) {
    const long long t2 = ((long long)q.b << (2 * L.lm)) + (q.i << L.lm) + q.j;
    return L.V[t2] * inv_dt + L.off[t2];
}

// Block partial sums of NP per-thread values into partial[s][p][blockIdx.x].
template <int NP>
__device__ __forceinline__ void block_partials(const double (&v)[NP], double *__restrict__ partial, int s) {
    __shared__ double red[NP][4];
#pragma unroll
    for (int p = 0; p < NP; p++) {
        double a = v[p];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if ((threadIdx.x & 63) == 0) red[p][threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x < NP) {
        const int p = threadIdx.x;
        partial[((long long)s * 3 + p) * gridDim.x + blockIdx.x] = ((red[p][0] + red[p][1]) + red[p][2]) + red[p][3];
    }
}

// Workgroup -> block-group order: the dispatcher deals workgroups round-robin
// to the 8 XCDs (each with its own L2); remapping gives every XCD a contiguous
// run of block groups, so face neighbours in phi and theta (adjacent block
// ids) are gathered from the same L2 while it still holds them.
__device__ __forceinline__ int xcd_group(int w, int nwg) {
    const int q = nwg >> 3, r = nwg & 7, x = w & 7;
    return x * q + (x < r ? x : r) + (w >> 3);
}

// ---- block-tile stencil kernels
// A workgroup of 256 threads owns 256/m^2 whole blocks (one for m = 16);
// thread (block, j, k) owns the column i = 0..m-1 (lanes of one i-plane are
// consecutive, so every load instruction is coalesced).  Phase 1 issues every
// load at once: the 6 ghost-table entries per thread, the x column, the op's
// input columns and the workgroup's coefficient tables.  Phase 2 gathers the
// face ghosts (1536 per workgroup for any m: the first donor of every entry,
// then the rest of the weighted ones) and stores the x tile (plane-major:
// tile[i][thread]), the ghosts and the coefficients in LDS.  Phase 3 applies
// the operator from registers (i +- 1) and LDS (j, k +- 1, ghosts, coefficients).

// trilinear inter/extrapolation inside the donor block from base cell c, in
// the term order of ModMath.f90:283-291 (weights may lie in [-0.5, 1.5]); f0 = X(c)
__device__ __forceinline__ double trilinear(const HLev &L, const Field &X, long long c, int wid, double f0) {
    const double wi = L.gw[3 * wid], wj = L.gw[3 * wid + 1], wk = L.gw[3 * wid + 2];
    const long long S1 = L.m, SI = (long long)L.m * L.m;
    const double f1 = X(c + SI), f2 = X(c + S1), f3 = X(c + 1), f4 = X(c + SI + S1), f5 = X(c + S1 + 1),
                 f6 = X(c + SI + 1), f7 = X(c + SI + S1 + 1);
    double t = f0 * (1.0 - wi) * (1.0 - wj) * (1.0 - wk);
    t = t + f1 * (wi) * (1.0 - wj) * (1.0 - wk);
    t = t + f2 * (1.0 - wi) * (wj) * (1.0 - wk);
    t = t + f3 * (1.0 - wi) * (1.0 - wj) * (wk);
    t = t + f4 * (wi) * (wj) * (1.0 - wk);
    t = t + f5 * (1.0 - wi) * (wj) * (wk);
    t = t + f6 * (wi) * (1.0 - wj) * (wk);
    t = t + f7 * (wi) * (wj) * (wk);
    return t;
}

// Per-cell epilogues: (io, x, A x, diag, in0, in1) -> stores and NP dot terms.
// IN0 / IN1: whether the op reads its input vectors in0 / in1 at the cell.
template <bool DOT>
struct OpApply {                 // y = A x; DOT: x.y, x.r
    double *y;
    const double *in0, *in1;     // r (DOT); unused
    static constexpr int NP = 2;
    static constexpr bool IN0 = DOT, IN1 = false;
    __device__ __forceinline__ void cell(long long io, double x, double ax, double, double r, double,
                                         double (&v)[NP]) const {
        y[io] = ax;
        if (DOT) { v[0] += x * ax; v[1] += x * r; }
    }
};
template <bool DOT>
struct OpJacobi {                // y = x + omega (r - A x)/diag; DOT: y.dq
    double *y;
    const double *in0, *in1;     // r; dq (DOT)
    double omega;
    static constexpr int NP = 1;
    static constexpr bool IN0 = true, IN1 = DOT;
    __device__ __forceinline__ void cell(long long io, double x, double ax, double diag, double r, double dq,
                                         double (&v)[NP]) const {
        const double yn = x + omega * (r - ax) / diag;
        y[io] = yn;
        if (DOT) v[0] += yn * dq;
    }
};
struct OpFeat {                  // r' = r - A z, u = r'/diag (float), u.u
    double *rp;
    float *u;
    const double *in0, *in1;     // r; unused
    static constexpr int NP = 1;
    static constexpr bool IN0 = true, IN1 = false;
    __device__ __forceinline__ void cell(long long io, double, double ax, double diag, double r, double,
                                         double (&v)[NP]) const {
        const double res = r - ax;
        rp[io] = res;
        const double w = res / diag;
        u[io] = (float)w;
        v[0] += w * w;
    }
};
struct OpApplyDot {              // c.r', c.Ac, c.q (Ac itself is not stored)
    const double *in0, *in1;     // r'; q
    static constexpr int NP = 3;
    static constexpr bool IN0 = true, IN1 = true;
    __device__ __forceinline__ void cell(long long, double x, double ax, double, double rp, double q,
                                         double (&v)[NP]) const {
        v[0] += x * rp;
        v[1] += x * ax;
        v[2] += x * q;
    }
};

// LDS layout (doubles): edge rows of the x tile (m = 16 only: the first and
// last j-row of every wavefront, [wave][2][i][k]), face ghosts [6 * 256],
// coefficient tables Ai [bpw][m+1][m], Aj [bpw][m][m+1], Ak, V, off [bpw][m][m]
template <int LM>
struct TileLds {
    static constexpr int m = 1 << LM, mm = m * m, bpw = 256 >> (2 * LM);
    static constexpr int NE = LM == 4 ? 4 * 2 * 16 * 16 : 0;
    static constexpr int G0 = NE, CI = G0 + 6 * 256, NI = bpw * (m + 1) * m, CJ = CI + NI, NJ = bpw * m * (m + 1),
                         CK = CJ + NJ, CV = CK + bpw * mm, CO = CV + bpw * mm, SIZE = CO + bpw * mm;
    static constexpr int NC = SIZE - CI;              // coefficient doubles
    static constexpr int NCT = (NC + 255) / 256;      // per thread
};

// x of lane l - 1 / l + 1 within rows of 16 lanes (DPP row shifts)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u & 0xffffffffull), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// phases 1-2: col = the thread's x column, a0 / a1 its columns of in0 / in1,
// sm = tile + ghosts + coefficients.  Ends with a barrier.
template <int LM, bool IN0, bool IN1>
__device__ __forceinline__ void load_tile(const HLev &L, const Field &F, int b0, bool own, long long cb,
                                          const double *__restrict__ in0, const double *__restrict__ in1, long long so,
                                          double (&col)[1 << LM], double (&a0)[1 << LM], double (&a1)[1 << LM],
                                          double *__restrict__ sm) {
    using T = TileLds<LM>;
    constexpr int m = T::m, mm = T::mm, bpw = T::bpw;
    const int tid = threadIdx.x;
    const int nbw = L.nb - b0 < bpw ? L.nb - b0 : bpw;
    const int ng = nbw * 6 * mm;
    const int2 *E = L.gsw + (long long)b0 * 6 * mm;
    int2 e[6];
#pragma unroll
    for (int u = 0; u < 6; u++) {
        const int g = tid + 256 * u;
        e[u] = g < ng ? E[g] : make_int2(-1, -1);
    }
#pragma unroll
    for (int i = 0; i < m; i++) {
        const long long c = cb + (long long)i * mm;
        col[i] = own ? F(c) : 0.0;
        if (IN0) a0[i] = own ? in0[so + c] : 0.0;
        if (IN1) a1[i] = own ? in1[so + c] : 0.0;
    }
    // coefficient tables of the workgroup's blocks: 5 contiguous ranges
    double cf[T::NCT];
#pragma unroll
    for (int u = 0; u < T::NCT; u++) {
        const int q = tid + 256 * u;
        double v = 0.0;
        if (q < T::NI) {
            if (q < nbw * (m + 1) * m) v = L.Ai[(long long)b0 * (m + 1) * m + q];
        } else if (q < T::NI + T::NJ) {
            const int r = q - T::NI;
            if (r < nbw * m * (m + 1)) v = L.Aj[(long long)b0 * m * (m + 1) + r];
        } else if (q < T::NC) {
            const int r = q - T::NI - T::NJ, which = r / (bpw * mm), w = r - which * (bpw * mm);
            const double *src = which == 0 ? L.Ak : (which == 1 ? L.V : L.off);
            if (w < nbw * mm) v = src[((long long)b0 << (2 * LM)) + w];
        }
        cf[u] = v;
    }
    // ghosts: the first donor value of every entry, then the weighted ones
    double f0[6];
#pragma unroll
    for (int u = 0; u < 6; u++) f0[u] = e[u].x >= 0 ? F(e[u].x) : 0.0;
#pragma unroll
    for (int u = 0; u < 6; u++) {
        double g = f0[u];
        if (e[u].y >= 0) g = trilinear(L, F, e[u].x, e[u].y, f0[u]);
        sm[T::G0 + tid + 256 * u] = g;
    }
#pragma unroll
    for (int u = 0; u < T::NCT; u++)
        if (tid + 256 * u < T::NC) sm[T::CI + tid + 256 * u] = cf[u];
    if (LM == 4) {
        // wavefront edge rows: lanes 0-15 (j = 4w) and 48-63 (j = 4w + 3)
        const int w = tid >> 6, l = tid & 63;
        if (l < 16 || l >= 48) {
            const int side = l < 16 ? 0 : 1;
#pragma unroll
            for (int i = 0; i < m; i++) sm[((w * 2 + side) * 16 + i) * 16 + (l & 15)] = col[i];
        }
    }
    __syncthreads();
}

// A x at cell i of the thread's column: i +- 1 from col, k +- 1 from the
// neighbouring lanes (DPP), j +- 1 from the lanes m apart (within the
// wavefront; for m = 16 the wavefront's first / last row reads the LDS edge
// rows of the neighbouring wavefront), face ghosts and coefficients from LDS.
// Every lane of the workgroup calls it (cross-lane operations).
template <int LM>
__device__ __forceinline__ double tile_apply(const double *__restrict__ sm, const double (&col)[1 << LM], int i, int lb,
                                             int j, int k, double inv_dt, double &diag) {
    using T = TileLds<LM>;
    constexpr int m = T::m, mm = T::mm;
    const int t = threadIdx.x, lane = t & 63;
    const int g = T::G0 + lb * 6 * mm;
    const double x0 = col[i];
    const double dkm = dpp_d<0x111>(x0), dkp = dpp_d<0x101>(x0);
    double djm = __shfl(x0, (lane - m) & 63), djp = __shfl(x0, (lane + m) & 63);
    if (LM == 4) {
        const int w = t >> 6;
        if (lane < 16 && w > 0) djm = sm[(((w - 1) * 2 + 1) * 16 + i) * 16 + k];
        if (lane >= 48 && w < 3) djp = sm[(((w + 1) * 2 + 0) * 16 + i) * 16 + k];
    }
    const double xim = i > 0 ? col[i > 0 ? i - 1 : 0] : sm[g + j * m + k];
    const double xip = i < m - 1 ? col[i < m - 1 ? i + 1 : 0] : sm[g + mm + j * m + k];
    const double xjm = j > 0 ? djm : sm[g + 2 * mm + i * m + k];
    const double xjp = j < m - 1 ? djp : sm[g + 3 * mm + i * m + k];
    const double xkm = k > 0 ? dkm : sm[g + 4 * mm + i * m + j];
    const double xkp = k < m - 1 ? dkp : sm[g + 5 * mm + i * m + j];
    const double *ai = sm + T::CI + lb * (m + 1) * m;
    const double *aj = sm + T::CJ + lb * m * (m + 1);
    const int t2 = lb * mm + i * m + j;
    const double cim = ai[i * m + j], cip = ai[(i + 1) * m + j];
    const double cjm = aj[i * (m + 1) + j], cjp = aj[i * (m + 1) + j + 1];
    const double ck = sm[T::CK + t2];
    diag = sm[T::CV + t2] * inv_dt + sm[T::CO + t2];
    double nb = cim * xim + cip * xip;
    nb = nb + cjm * xjm + cjp * xjp;
    nb = nb + ck * xkm + ck * xkp;
    return diag * x0 - nb;
}

// One operator pass over every cell of a level: op.cell at each cell, dot
// products per workgroup into partial (when given).
template <int LM, class OP>
__global__ void __launch_bounds__(256) k_hmarch(HLev L, const double *__restrict__ x, const double *__restrict__ inv_dt,
                                                const int *__restrict__ active, long long istride, OP op,
                                                double *__restrict__ partial) {
    constexpr int m = 1 << LM, mm = m * m, bpw = 256 >> (2 * LM);
    __shared__ double sm[TileLds<LM>::SIZE];
    const int s = blockIdx.y;
    if (!active[s]) return;
    const int tid = threadIdx.x;
    const int lb = tid >> (2 * LM), j = (tid >> LM) & (m - 1), k = tid & (m - 1);
    const int b0 = xcd_group(blockIdx.x, gridDim.x) * bpw, b = b0 + lb;
    const bool own = b < L.nb;
    const Field F{x + s * istride};
    const long long so = s * istride;
    const long long cb = ((long long)b << (3 * LM)) + (j << LM) + k;
    double col[m], a0[m], a1[m];
    load_tile<LM, OP::IN0, OP::IN1>(L, F, b0, own, cb, op.in0, op.in1, so, col, a0, a1, sm);
    double v[OP::NP] = {};
    const double idt = inv_dt[s];
#pragma unroll
    for (int i = 0; i < m; i++) {
        double diag;
        const double ax = tile_apply<LM>(sm, col, i, lb, j, k, idt, diag);
        if (own) op.cell(so + cb + (long long)i * mm, col[i], ax, diag, OP::IN0 ? a0[i] : 0.0, OP::IN1 ? a1[i] : 0.0, v);
    }
    if (partial) block_partials<OP::NP>(v, partial, s);
}

// residual restriction fused with the residual: coarse cell = sum over its 8
// children (di, dj, dk order) of r - A x (integrated residual).  The children
// columns (j, k), (j, k+1), (j+1, k), (j+1, k+1) of an even (j, k) lie in one
// wavefront (lanes t, t + 1, t + m, t + m + 1): their residuals are gathered
// by lane shuffles, the i pair from the thread's own column.
template <int LM>
__global__ void __launch_bounds__(256) k_hresrestrict(HLev L, const double *__restrict__ x, const double *__restrict__ r,
                                                      double *__restrict__ rc, const double *__restrict__ inv_dt,
                                                      const int *__restrict__ active, long long istride,
                                                      long long cstride) {
    constexpr int m = 1 << LM, mm = m * m, bpw = 256 >> (2 * LM), LC = LM - 1;
    __shared__ double sm[TileLds<LM>::SIZE];
    const int s = blockIdx.y;
    if (!active[s]) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int lb = tid >> (2 * LM), j = (tid >> LM) & (m - 1), k = tid & (m - 1);
    const int b0 = xcd_group(blockIdx.x, gridDim.x) * bpw, b = b0 + lb;
    const bool own = b < L.nb;
    const Field F{x + s * istride};
    const long long so = s * istride;
    const long long cb = ((long long)b << (3 * LM)) + (j << LM) + k;
    double col[m], a0[m], a1[m];
    load_tile<LM, true, false>(L, F, b0, own, cb, r, nullptr, so, col, a0, a1, sm);
    const double idt = inv_dt[s];
    const bool out = own && !(j & 1) && !(k & 1);
    double *R = rc + s * cstride + ((long long)b << (3 * LC)) + ((j >> 1) << LC) + (k >> 1);
#pragma unroll
    for (int ic = 0; ic < m / 2; ic++) {
        double diag;
        const double ax0 = tile_apply<LM>(sm, col, 2 * ic, lb, j, k, idt, diag);
        const double ax1 = tile_apply<LM>(sm, col, 2 * ic + 1, lb, j, k, idt, diag);
        const double o0 = own ? a0[2 * ic] - ax0 : 0.0, o1 = own ? a0[2 * ic + 1] - ax1 : 0.0;
        const double k0 = __shfl(o0, (lane + 1) & 63, 64), k1 = __shfl(o1, (lane + 1) & 63, 64);
        const double j0 = __shfl(o0, (lane + m) & 63, 64), j1 = __shfl(o1, (lane + m) & 63, 64);
        const double q0 = __shfl(o0, (lane + m + 1) & 63, 64), q1 = __shfl(o1, (lane + m + 1) & 63, 64);
        double t = 0.0;
        t += o0; t += k0; t += j0; t += q0;
        t += o1; t += k1; t += j1; t += q1;
        if (out) R[(long long)ic << (2 * LC)] = t;
    }
}

// Jacobi from x = 0: y = (omega r)/diag (the damped Jacobi value at x = 0)
__global__ void __launch_bounds__(256) k_hjacobi0(HLev L, const double *__restrict__ r, double *__restrict__ y,
                                                  const double *__restrict__ inv_dt, const int *__restrict__ active,
                                                  long long istride, double omega) {
    const int s = blockIdx.y;
    if (!active[s]) return;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)L.nb * L.M3) return;
    const long long io = s * istride + c;
    y[io] = (omega * r[io]) / diag_at(L, cell_of(L, c), inv_dt[s]);
}

// Linear prolongation, stored: fine cell (i, j, k) of level LM adds
//   z0 + (zi - z0)/4 + (zj - z0)/4 + (zk - z0)/4
// with z0 its parent on the coarse level and zi, zj, zk the parent's face
// neighbours on the fine cell's side in each axis (cell-centred linear
// interpolation from face neighbours only, exact for fields linear in the
// block's index space).  A neighbour outside the coarse block is the coarse
// level's face ghost (copy or trilinear from the donor table); at an insulated
// radial face (no donor) it is z0 (zero gradient).
template <int LM>
__global__ void __launch_bounds__(256) k_hprolong_lin(HLev Lc, double *__restrict__ zf, const double *__restrict__ zc,
                                                      const int *__restrict__ active, long long fstride,
                                                      long long cstride) {
    constexpr int m = 1 << LM, LC = LM - 1, mc = 1 << LC, mmc = mc * mc;
    const int s = blockIdx.y;
    if (!active[s]) return;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ((long long)Lc.nb << (3 * LM))) return;
    const int b = (int)(c >> (3 * LM));
    const int q = (int)(c & ((1 << (3 * LM)) - 1));
    const int i = q >> (2 * LM), j = (q >> LM) & (m - 1), k = q & (m - 1);
    const int ic = i >> 1, jc = j >> 1, kc = k >> 1;
    const int si = (i & 1) ? 1 : -1, sj = (j & 1) ? 1 : -1, sk = (k & 1) ? 1 : -1;
    const double *Z = zc + s * cstride;
    const Field F{Z};
    const long long cb = ((long long)b << (3 * LC)) + (ic << (2 * LC)) + (jc << LC) + kc;
    const double z0 = Z[cb];
    const int2 *E = Lc.gsw + (long long)b * 6 * mmc;
    auto ghost = [&](int face, int idx) -> double {
        const int2 e = E[face * mmc + idx];
        if (e.x < 0) return z0;
        const double f0 = Z[e.x];
        return e.y < 0 ? f0 : trilinear(Lc, F, e.x, e.y, f0);
    };
    const int in = ic + si, jn = jc + sj, kn = kc + sk;
    const double zi = (in >= 0 && in < mc) ? Z[cb + si * mmc] : ghost(si < 0 ? 0 : 1, jc * mc + kc);
    const double zj = (jn >= 0 && jn < mc) ? Z[cb + sj * mc] : ghost(sj < 0 ? 2 : 3, ic * mc + kc);
    const double zk = (kn >= 0 && kn < mc) ? Z[cb + sk] : ghost(sk < 0 ? 4 : 5, ic * mc + jc);
    double v = z0 + 0.25 * (zi - z0);
    v = v + 0.25 * (zj - z0);
    v = v + 0.25 * (zk - z0);
    const long long io = s * fstride + c;
    zf[io] = zf[io] + v;
}

// Sum over one workgroup of 1024 threads: every wavefront reduces by halving
// shifts, then thread 0 adds the 16 wavefront sums in order; all threads get
// the result.
__device__ __forceinline__ double wg1024_sum(double a, double *red) {
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int w = 1; w < 16; w++) t = t + red[w];
        red[16] = t;
    }
    __syncthreads();
    return red[16];
}

// Coarsest level (one cell per block): Jacobi-preconditioned CG from x = 0,
// ModLinearSolver_JCB_CG (3_5/ModLinearSolver.f90:86-180) as a method, with
// the Jacobi diagonal as preconditioner and a relative stop r.z <= tol2 (r.z)_0
// or max_iter iterations.  One workgroup per system holds every row
// (thread t: blocks t, t + 1024, ...); r, p and A p live in work[s][3][nb].
// A p per row is the level's operator with face ghosts copied from the donor
// blocks' single cells (0 at insulated faces, whose coefficient is 0).
__global__ void __launch_bounds__(1024) k_hcoarse_cg(HLev L, const double *__restrict__ rhs, double *__restrict__ x,
                                                     double *__restrict__ work, const double *__restrict__ inv_dt,
                                                     const int *__restrict__ active, long long stride, double tol2,
                                                     int max_iter) {
    __shared__ double red[17];
    const int s = blockIdx.x;
    if (!active[s]) return;
    const int nb = L.nb, t = threadIdx.x;
    const double idt = inv_dt[s];
    double *X = x + s * stride;
    const double *B = rhs + s * stride;
    double *R = work + (long long)s * 3 * nb, *Pv = R + nb, *AP = Pv + nb;
    double acc = 0.0;
    for (int b = t; b < nb; b += 1024) {
        const double dg = L.V[b] * idt + L.off[b];
        const double r = B[b];
        R[b] = r;
        const double z = r / dg;
        Pv[b] = z;
        X[b] = 0.0;
        acc += r * z;
    }
    double rz = wg1024_sum(acc, red);
    const double stop = tol2 * rz;
    for (int it = 0; it < max_iter && rz > stop; it++) {
        acc = 0.0;
        for (int b = t; b < nb; b += 1024) {
            const int2 *E = L.gsw + (long long)b * 6;
            double g[6];
#pragma unroll
            for (int f = 0; f < 6; f++) {
                const int2 e = E[f];
                g[f] = e.x >= 0 ? Pv[e.x] : 0.0;
            }
            const double cim = L.Ai[2 * b], cip = L.Ai[2 * b + 1], cjm = L.Aj[2 * b], cjp = L.Aj[2 * b + 1];
            const double ck = L.Ak[b];
            const double diag = L.V[b] * idt + L.off[b];
            double nbv = cim * g[0] + cip * g[1];
            nbv = nbv + cjm * g[2] + cjp * g[3];
            nbv = nbv + ck * g[4] + ck * g[5];
            const double ap = diag * Pv[b] - nbv;
            AP[b] = ap;
            acc += Pv[b] * ap;
        }
        const double pAp = wg1024_sum(acc, red);
        if (!(pAp > 0.0)) break;
        const double alpha = rz / pAp;
        acc = 0.0;
        for (int b = t; b < nb; b += 1024) {
            const double dg = L.V[b] * idt + L.off[b];
            X[b] = X[b] + alpha * Pv[b];
            const double r = R[b] - alpha * AP[b];
            R[b] = r;
            acc += r * (r / dg);
        }
        const double rzn = wg1024_sum(acc, red);
        const double beta = rzn / rz;
        rz = rzn;
        for (int b = t; b < nb; b += 1024) {
            const double dg = L.V[b] * idt + L.off[b];
            Pv[b] = R[b] / dg + beta * Pv[b];
        }
        __syncthreads();   // every p before the next product reads its neighbours
    }
}

__global__ void __launch_bounds__(256) k_hzero(double *__restrict__ a, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// ---- elementwise level-0 kernels: a workgroup owns kElem consecutive cells,
// thread t the cells t, t + 256, ...; dot products per workgroup
constexpr int kElem = 4096;

// up to 2 dot products per system: partial[s][p][wg]
__global__ void __launch_bounds__(256) k_hdot(long long n, const double *__restrict__ a0, const double *__restrict__ b0,
                                              const double *__restrict__ a1, const double *__restrict__ b1, int npairs,
                                              const int *__restrict__ active, double *__restrict__ partial,
                                              long long istride) {
    const int s = blockIdx.y;
    if (active && !active[s]) return;
    double v[2] = {0.0, 0.0};
    const long long c0 = (long long)blockIdx.x * kElem + threadIdx.x;
#pragma unroll 4
    for (int u = 0; u < kElem / 256; u++) {
        const long long c = c0 + 256 * u;
        if (c < n) {
            const long long io = s * istride + c;
            v[0] += a0[io] * b0[io];
            if (npairs > 1) v[1] += a1[io] * b1[io];
        }
    }
    block_partials<2>(v, partial, s);
}

// sum of nwg partials per (system, pair), fixed order
__global__ void __launch_bounds__(256) k_hdot_final(const double *__restrict__ partial, int nwg, int npairs,
                                                   double *__restrict__ out) {
    __shared__ double red[4];
    const int s = blockIdx.x;
    for (int p = 0; p < npairs; p++) {
        const double *P = partial + ((long long)s * 3 + p) * nwg;
        double v0 = 0.0, v1 = 0.0;
        int w = threadIdx.x;
        for (; w + 256 < nwg; w += 512) { v0 += P[w]; v1 += P[w + 256]; }
        if (w < nwg) v0 += P[w];
        double v = v0 + v1;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) out[s * 3 + p] = ((red[0] + red[1]) + red[2]) + red[3];
        __syncthreads();
    }
}

// Per-system scalars on the device (one thread per system), kSc slots each:
// [0] alpha = <p,r>/<p,q>, [1] beta = <z,q>/gamma, [2] gamma = <p,q>,
// [3] feature scale of the learned correction, [4] its step along c, [5] <b,b>,
// [6] <z,q> of the V-cycle output.
constexpr int kSc = 8;
// what 0: alpha and gamma (dots: [0]=<p,q>, [1]=<p,r>)
// what 1: beta (dots: [0]=<z,q>)
// what 2: correction step <c,r'>/<c,Ac> (dots: [0]=<c,r'>, [1]=<c,Ac>), 0 unless <c,Ac> > 0
// what 3: feature scale sqrt(sum / ncell) (dots: [0]=sum (r'/diag)^2)
// what 4: <b,b> (dots: [0])
// what 5: beta of the learned direction from [6], [4] and dots[2] = <c,q>
// what 6: [6] = <z,q> (dots: [0])
// what 7: beta of the learned direction with per-block steps (dots: [0] = sum_b a_b <c_b,q_b>)
__global__ void k_hscalars(int what, int nsys, const double *__restrict__ dots, double *__restrict__ sc,
                           const int *__restrict__ active, double ncell) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    double *S = sc + kSc * s;
    if (what == 0) {
        S[2] = dots[3 * s];
        S[0] = active[s] ? dots[3 * s + 1] / dots[3 * s] : 0.0;
    } else if (what == 1) {
        S[1] = active[s] ? dots[3 * s] / S[2] : 0.0;
    } else if (what == 2) {
        const double num = dots[3 * s], den = dots[3 * s + 1];
        const double a = num / den;
        S[4] = (active[s] && den > 0.0 && a == a && a < 1e300 && a > -1e300) ? a : 0.0;
    } else if (what == 3) {
        S[3] = active[s] ? sqrt(dots[3 * s] / ncell) : 0.0;
    } else if (what == 4) {
        S[5] = dots[3 * s];
    } else if (what == 7) {
        // beta of the learned direction with per-block steps: (<z,q> + sum_b a_b <c_b,q_b>)/gamma
        S[1] = active[s] ? (S[6] + dots[3 * s]) / S[2] : 0.0;
    } else if (what == 5) {
        // beta of the learned direction: <z + a c, q>/gamma = (<z,q> + a <c,q>)/gamma
        S[1] = active[s] ? (S[6] + S[4] * dots[3 * s + 2]) / S[2] : 0.0;
    } else {
        S[6] = dots[3 * s];                 // <z, q> of the V-cycle output
    }
}

// Convergence on the device (one workgroup): for every system still active,
// count the iteration, store ||r||/||b|| and deactivate it at tol (or on a
// non-finite residual); out[0] = number of systems left active.
__global__ void __launch_bounds__(256) k_hconverge(int nsys, const double *__restrict__ dots,
                                                   const double *__restrict__ sc, double tol, int *__restrict__ active,
                                                   int *__restrict__ iters, double *__restrict__ relres,
                                                   int *__restrict__ out) {
    __shared__ int cnt[256];
    int c = 0;
    for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
        if (!active[s]) continue;
        const double rel = sqrt(dots[3 * s] / sc[kSc * s + 5]);
        iters[s] = iters[s] + 1;
        relres[s] = rel;
        if (rel <= tol || !(rel == rel) || rel > 1e300) active[s] = 0;
        else c++;
    }
    cnt[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < (int)blockDim.x; k++) t += cnt[k];
        out[0] = t;
    }
}

__global__ void k_hizero(int *__restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// x += alpha p ; r -= alpha q, and partial sums of r.r
__global__ void __launch_bounds__(256) k_hupdate_xr(long long n, double *__restrict__ x, double *__restrict__ r,
                                                    const double *__restrict__ p, const double *__restrict__ q,
                                                    const double *__restrict__ sc, const int *__restrict__ active,
                                                    double *__restrict__ partial, long long istride) {
    const int s = blockIdx.y;
    if (!active[s]) return;
    const double a = sc[kSc * s];
    double v[1] = {0.0};
    const long long c0 = (long long)blockIdx.x * kElem + threadIdx.x;
#pragma unroll 4
    for (int u = 0; u < kElem / 256; u++) {
        const long long c = c0 + 256 * u;
        if (c < n) {
            const long long io = s * istride + c;
            x[io] = x[io] + a * p[io];
            const double rn = r[io] - a * q[io];
            r[io] = rn;
            v[0] += rn * rn;
        }
    }
    block_partials<1>(v, partial, s);
}

// p = z - beta p; beta = 0 starts a fresh direction
__global__ void __launch_bounds__(256) k_hupdate_p(long long n, double *__restrict__ p, const double *__restrict__ z,
                                                   const double *__restrict__ sc, const int *__restrict__ active,
                                                   long long istride, int use_beta) {
    const int s = blockIdx.y;
    if (!active[s]) return;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double be = use_beta ? sc[kSc * s + 1] : 0.0;
    const long long io = s * istride + c;
    p[io] = z[io] - be * p[io];
}

// b = (V/dt) T
__global__ void __launch_bounds__(256) k_hrhs(HLev L, const double *__restrict__ T, double *__restrict__ bvec,
                                              const double *__restrict__ inv_dt, long long istride) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)L.nb * L.M3) return;
    const int s = blockIdx.y;
    const Cell q = cell_of(L, c);
    const long long t2 = ((long long)q.b << (2 * L.lm)) + (q.i << L.lm) + q.j;
    bvec[s * istride + c] = L.V[t2] * inv_dt[s] * T[s * istride + c];
}

// PFSS field B = -grad(phi) of the finest level's solution x (system 0) at
// every cell: the second-order gradient of ModSpherical_Grad_f_O2
// (3_6/ModSpherical.f90:161-199: (1./2./dx)*(f(+1) - f(-1)), then 1/r and
// 1/sin(theta)) with the sign 3_5/ModPFSS.f90:141 states.  Neighbours across a
// block face are the operator's face ghosts; at the domain's radial boundaries
// the ghosts of the boundary conditions: bottom phi_1 + bot (= Br dr, the
// insulated face carrying the flux Br), top -phi_m (phi = 0 on the face).
// geo per block: 1/r (m), 1/sin(theta) (m), (1./2./dx) of r, theta, phi;
// B: [block][3][i][j][k].
__global__ void __launch_bounds__(256) k_pfss_field(HLev L, const double *__restrict__ x, const double *__restrict__ bot,
                                                    const double *__restrict__ geo, double *__restrict__ B) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)L.nb * L.M3) return;
    const Cell q = cell_of(L, c);
    const int m = L.m;
    const long long mm = (long long)m * m, cb = (long long)q.b * L.M3;
    const Field F{x};
    // the face ghost of face f (i-, i+, j-, j+, k-, k+) at in-face position (t1, t2)
    auto ghost = [&](int f, int t1, int t2) -> double {
        const int2 e = L.gsw[((long long)q.b * 6 + f) * mm + (long long)t1 * m + t2];
        if (e.x >= 0) {
            const double f0 = F(e.x);
            return e.y >= 0 ? trilinear(L, F, e.x, e.y, f0) : f0;
        }
        // no donor: a radial boundary face of the domain
        return f == 0 ? F(cb + ((long long)q.j * m + q.k)) + bot[(long long)q.b * mm + (long long)q.j * m + q.k]
                      : -F(cb + ((long long)(m - 1) * m + q.j) * m + q.k);
    };
    const double xim = q.i > 0 ? F(c - mm) : ghost(0, q.j, q.k);
    const double xip = q.i < m - 1 ? F(c + mm) : ghost(1, q.j, q.k);
    const double xjm = q.j > 0 ? F(c - m) : ghost(2, q.i, q.k);
    const double xjp = q.j < m - 1 ? F(c + m) : ghost(3, q.i, q.k);
    const double xkm = q.k > 0 ? F(c - 1) : ghost(4, q.i, q.j);
    const double xkp = q.k < m - 1 ? F(c + 1) : ghost(5, q.i, q.j);
    const double *G = geo + (long long)q.b * (2 * m + 3);
    const double rinv = G[q.i], sinv = G[m + q.j];
    const double gr = G[2 * m] * (xip - xim);
    double gt = G[2 * m + 1] * (xjp - xjm);
    gt = gt * rinv;
    double gp = G[2 * m + 2] * (xkp - xkm);
    gp = gp * rinv;
    gp = gp * sinv;
    const long long o = (long long)q.b * 3 * L.M3 + (c - cb);
    B[o] = -gr;
    B[o + L.M3] = -gt;
    B[o + 2 * L.M3] = -gp;
}

__global__ void __launch_bounds__(256) k_hcopy(const double *__restrict__ a, double *__restrict__ b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ---- learned correction of the V-cycle (features: k_hmarch with OpFeat;
// the step's dots: k_hmarch with OpApplyDot)
// p = z + a c - beta p: the learned direction (a = the correction step);
// beta = 0 starts a fresh direction
__global__ void __launch_bounds__(256) k_hupdate_p2(long long n, double *__restrict__ p, const double *__restrict__ z,
                                                    const double *__restrict__ cc, const double *__restrict__ sc,
                                                    const double *__restrict__ ablk, int nb, const int *__restrict__ active,
                                                    long long istride, int use_beta) {
    const int s = blockIdx.y;
    if (!active[s]) return;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double a = ablk ? ablk[(long long)s * nb + (c >> 12)] : sc[kSc * s + 4];
    const double be = use_beta ? sc[kSc * s + 1] : 0.0;
    const long long io = s * istride + c;
    p[io] = (z[io] + a * cc[io]) - be * p[io];
}

// Per-block correction steps (16^3 blocks, one block per workgroup of the
// correction's dot pass): a_b = <c_b,r'_b>/<c_b,(Ac)_b> (0 unless the
// denominator is positive), and in partial slot 0 the block's a_b <c_b,q_b>
__global__ void __launch_bounds__(256) k_hblockstep(int nwg, int nb, double *__restrict__ partial,
                                                    double *__restrict__ ablk, const int *__restrict__ active) {
    const int s = blockIdx.y;
    if (!active[s]) return;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwg) return;
    const int b = xcd_group(w, nwg);
    double *P = partial + (long long)s * 3 * nwg;
    const double num = P[w], den = P[nwg + w], cq = P[2 * nwg + w];
    const double a = num / den;
    const double ab = (den > 0.0 && a == a && a < 1e300 && a > -1e300) ? a : 0.0;
    ablk[(long long)s * nb + b] = ab;
    P[w] = ab * cq;
}

// training records: copy the network input image and w = x + z of record slot
// r (system rs[r], block rb[r]) into X[r] (float) and W[r] (double)
__global__ void __launch_bounds__(256) k_hrecord(HLev L, const float *__restrict__ xin, const double *__restrict__ x,
                                                 const double *__restrict__ z, const int *__restrict__ rs,
                                                 const int *__restrict__ rb, float *__restrict__ X, double *__restrict__ W,
                                                 const double *__restrict__ sc, double *__restrict__ rscale,
                                                 long long istride) {
    const long long r = blockIdx.x;
    const int s = rs[r], b = rb[r];
    if (threadIdx.x == 0) rscale[r] = sc[kSc * s + 3];
    for (int q = threadIdx.x; q < (int)L.M3; q += blockDim.x) {
        const long long io = s * istride + (long long)b * L.M3 + q;
        const double scale = sc[kSc * s + 3];
        X[r * L.M3 + q] = scale > 0.0 ? (float)((double)xin[io] / scale) : 0.0f;
        W[r * L.M3 + q] = x[io] + z[io];
    }
}

// targets: Y[r] = (x*[system][block] - W[r]) / scale[r]
__global__ void __launch_bounds__(256) k_htarget(HLev L, const double *__restrict__ x, const int *__restrict__ rs,
                                                 const int *__restrict__ rb, const double *__restrict__ rscale,
                                                 const double *__restrict__ W, float *__restrict__ Y, long long istride) {
    const long long r = blockIdx.x;
    const int s = rs[r], b = rb[r];
    const double sc = rscale[r];
    for (int q = threadIdx.x; q < (int)L.M3; q += blockDim.x) {
        const double e = x[s * istride + (long long)b * L.M3 + q] - W[r * L.M3 + q];
        Y[r * L.M3 + q] = sc > 0.0 ? (float)(e / sc) : 0.0f;
    }
}

// ------------------------------------------------------------------ solver

struct HeatSolver::Impl {
    int nsys = 1, nb = 0, nlev = 0;
    hipStream_t stream = nullptr;
    std::vector<HeatLevel> host;
    std::vector<HLev> L;
    std::vector<void *> allocs;
    // per level work arrays (device, interior cells): smoother iterate pair, right-hand side
    std::vector<double *> Z, Z2, R;
    // level-0 FCG arrays
    double *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *b = nullptr, *tmp = nullptr;
    double *inv_dt = nullptr, *partial = nullptr, *dots = nullptr, *sc = nullptr;
    int *active = nullptr;
    // learned correction work arrays (allocated on first use)
    double *rp = nullptr, *cc = nullptr;
    float *xin = nullptr;
    int learned_pre0 = 2, learned_post0 = 2;
    bool blockstep = true;           // per-block correction steps
    double *ablk = nullptr;
    // captured FCG iteration per mode (plain, learned) and device-side convergence state
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    double graph_tol[2] = {0.0, 0.0};   // tolerance baked into each captured k_hconverge
    bool use_graph = true;
    int learned_iters = 0;              // directions built with the learned correction (0: all)
    int *iters_dev = nullptr, *nactive_dev = nullptr;
    double *relres_dev = nullptr;
    void drop_graphs() {
        for (auto &g : graph)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    }
    // workgroups of the level-0 kernels: block-marching (256/m^2 blocks each)
    // and elementwise (kElem cells each); their dot partials are per workgroup
    unsigned march_wgs(int l) const {
        const int bpw = 256 >> (2 * L[l].lm);
        return (unsigned)((nb + bpw - 1) / bpw);
    }
    unsigned elem_wgs() const { return (unsigned)((ist(0) + kElem - 1) / kElem); }
    dim3 grid_elem() const { return dim3(elem_wgs(), (unsigned)nsys); }
    void ensure_learned();
    void features(double *z);
    void learned_correction(bool first);
    void final_dots(int npairs, unsigned nwg) {
        hipLaunchKernelGGL(k_hdot_final, dim3((unsigned)nsys), dim3(256), 0, stream, (const double *)partial, (int)nwg,
                           npairs, dots);
    }
    void scalars(int what, double ncell = 0.0) {
        hipLaunchKernelGGL(k_hscalars, dim3((nsys + 63) / 64), dim3(64), 0, stream, what, nsys, (const double *)dots, sc,
                           (const int *)active, ncell);
    }
    int pre = 2, post = 2;
    // coarsest level: Jacobi-preconditioned CG to r.z <= coarse_tol^2 (r.z)_0
    double coarse_tol = 1.0e-2;
    int coarse_max = 50;
    double *coarse_work = nullptr;   // [system][r, p, Ap][nb]
    double omega = 0.8;
    std::vector<double> dt;
    double dt_expl = 0.0;
    HeatCorrector *corr = nullptr;

    double *dalloc(size_t n) {
        void *p_ = nullptr;
        HCHECK(hipMalloc(&p_, std::max<size_t>(n, 1) * sizeof(double)));
        HCHECK(hipMemset(p_, 0, std::max<size_t>(n, 1) * sizeof(double)));
        allocs.push_back(p_);
        return (double *)p_;
    }
    template <typename T>
    T *upload(const std::vector<T> &v) {
        void *p_ = nullptr;
        HCHECK(hipMalloc(&p_, std::max<size_t>(v.size(), 1) * sizeof(T)));
        if (!v.empty()) HCHECK(hipMemcpy(p_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        allocs.push_back(p_);
        return (T *)p_;
    }
    ~Impl() {
        drop_graphs();
        for (void *a : allocs) (void)hipFree(a);
        if (stream) (void)hipStreamDestroy(stream);
    }
    long long ist(int l) const { return (long long)nb * L[l].M3; }
    dim3 grid_cells(int l) const { return dim3((unsigned)((ist(l) + 255) / 256), (unsigned)nsys); }
    void zero(double *a, long long n) {
        hipLaunchKernelGGL(k_hzero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, n);
    }
    void set_all_active() {
        std::vector<int> one(nsys, 1);
        HCHECK(hipMemcpy(active, one.data(), sizeof(int) * nsys, hipMemcpyHostToDevice));
    }
    template <class OP>
    void launch_march(int l, const double *x_, const OP &op, double *part) {
        const dim3 g(march_wgs(l), (unsigned)nsys);
        const HLev &D = L[l];
        const double *idt = inv_dt;
        const int *ac = active;
        const long long is = ist(l);
        switch (D.lm) {
            case 0: hipLaunchKernelGGL((k_hmarch<0, OP>), g, dim3(256), 0, stream, D, x_, idt, ac, is, op, part); break;
            case 1: hipLaunchKernelGGL((k_hmarch<1, OP>), g, dim3(256), 0, stream, D, x_, idt, ac, is, op, part); break;
            case 2: hipLaunchKernelGGL((k_hmarch<2, OP>), g, dim3(256), 0, stream, D, x_, idt, ac, is, op, part); break;
            case 3: hipLaunchKernelGGL((k_hmarch<3, OP>), g, dim3(256), 0, stream, D, x_, idt, ac, is, op, part); break;
            case 4: hipLaunchKernelGGL((k_hmarch<4, OP>), g, dim3(256), 0, stream, D, x_, idt, ac, is, op, part); break;
            default: throw std::runtime_error("heat blocks are at most 16^3");
        }
    }
    // R[l+1] = restriction of R[l] - A Z[l]
    void restrict_residual(int l) {
        const dim3 g(march_wgs(l), (unsigned)nsys);
        const HLev &D = L[l];
        const double *z = Z[l], *rr = R[l], *idt = inv_dt;
        const int *ac = active;
        switch (D.lm) {
            case 1: hipLaunchKernelGGL(k_hresrestrict<1>, g, dim3(256), 0, stream, D, z, rr, R[l + 1], idt, ac, ist(l), ist(l + 1)); break;
            case 2: hipLaunchKernelGGL(k_hresrestrict<2>, g, dim3(256), 0, stream, D, z, rr, R[l + 1], idt, ac, ist(l), ist(l + 1)); break;
            case 3: hipLaunchKernelGGL(k_hresrestrict<3>, g, dim3(256), 0, stream, D, z, rr, R[l + 1], idt, ac, ist(l), ist(l + 1)); break;
            case 4: hipLaunchKernelGGL(k_hresrestrict<4>, g, dim3(256), 0, stream, D, z, rr, R[l + 1], idt, ac, ist(l), ist(l + 1)); break;
            default: throw std::runtime_error("no restriction below 2^3 cells per block");
        }
    }
    // y = A x (and per-workgroup partial sums of x.y, x.r when r is given)
    void apply0(const double *x_, double *y, const double *r_dot) {
        if (r_dot) launch_march(0, x_, OpApply<true>{y, r_dot, nullptr}, partial);
        else launch_march(0, x_, OpApply<false>{y, nullptr, nullptr}, nullptr);
    }
    // one damped Jacobi sweep Z[l] -> Z2[l] (swapped); dq: accumulate
    // <new z, dq> (level 0)
    void sweep(int l, const double *dq) {
        if (dq) launch_march(l, Z[l], OpJacobi<true>{Z2[l], R[l], dq, omega}, partial);
        else launch_march(l, Z[l], OpJacobi<false>{Z2[l], R[l], nullptr, omega}, nullptr);
        std::swap(Z[l], Z2[l]);
    }
    // first sweep from zero: Z[l] = omega R[l] / diag
    void sweep0(int l) {
        hipLaunchKernelGGL(k_hjacobi0, grid_cells(l), dim3(256), 0, stream, L[l], (const double *)R[l], Z[l],
                           (const double *)inv_dt, (const int *)active, ist(l), omega);
    }

    // zf (level l) += linear prolongation of level l+1
    void prolong(int l) {
        const HLev &Lc = L[l + 1];
        const dim3 g = grid_cells(l);
        double *zf = Z[l];
        const double *zc = Z[l + 1];
        const int *ac = active;
        const long long fs = ist(l), cs = ist(l + 1);
        switch (L[l].lm) {
            case 1: hipLaunchKernelGGL(k_hprolong_lin<1>, g, dim3(256), 0, stream, Lc, zf, zc, ac, fs, cs); break;
            case 2: hipLaunchKernelGGL(k_hprolong_lin<2>, g, dim3(256), 0, stream, Lc, zf, zc, ac, fs, cs); break;
            case 3: hipLaunchKernelGGL(k_hprolong_lin<3>, g, dim3(256), 0, stream, Lc, zf, zc, ac, fs, cs); break;
            case 4: hipLaunchKernelGGL(k_hprolong_lin<4>, g, dim3(256), 0, stream, Lc, zf, zc, ac, fs, cs); break;
            default: throw std::runtime_error("no prolongation onto 1-cell blocks");
        }
    }
    // coarsest level: Z[l] = CG solve of its operator with right-hand side R[l]
    void coarse_solve(int l) {
        hipLaunchKernelGGL(k_hcoarse_cg, dim3((unsigned)nsys), dim3(1024), 0, stream, L[l], (const double *)R[l], Z[l],
                           coarse_work, (const double *)inv_dt, (const int *)active, ist(l),
                           coarse_tol * coarse_tol, coarse_max);
    }

    // V-cycle: z0 (level 0) = M r0.  pre0 / post0: smoothing sweeps on the
    // finest level (others use pre / post); the coarsest level is solved by
    // CG.  dq: accumulate <z0, dq> into the partial sums; returns the number of
    // workgroup partials to sum (0 without dq).
    unsigned vcycle(const double *r0, double *&z0, int pre0, int post0, const double *dq) {
        R[0] = const_cast<double *>(r0);
        for (int l = 0; l < nlev; l++) {
            if (l == nlev - 1) {
                coarse_solve(l);
                break;
            }
            const int sweeps = l == 0 ? pre0 : pre;
            if (sweeps == 0) {
                zero(Z[l], ist(l) * nsys);
            } else {
                sweep0(l);
                for (int s = 1; s < sweeps; s++) sweep(l, nullptr);
            }
            restrict_residual(l);
        }
        unsigned nwg = 0;
        for (int l = nlev - 2; l >= 0; l--) {
            prolong(l);
            const int sweeps = l == 0 ? post0 : post;
            for (int s = 0; s < sweeps; s++) {
                const bool last0 = l == 0 && s == sweeps - 1 && dq;
                sweep(l, last0 ? dq : nullptr);
                if (last0) nwg = march_wgs(0);
            }
        }
        if (dq && !nwg) {
            hipLaunchKernelGGL(k_hdot, grid_elem(), dim3(256), 0, stream, ist(0), (const double *)Z[0], dq,
                               (const double *)nullptr, (const double *)nullptr, 1, (const int *)active, partial, ist(0));
            nwg = elem_wgs();
        }
        z0 = Z[0];
        return nwg;
    }
};

HeatSolver::HeatSolver() : d(new Impl) {}
HeatSolver::~HeatSolver() { delete d; }

void HeatSolver::setup(const Layout &lay, const HeatParams &hp, int nsys, int nthreads) {
    Impl &I = *d;
    I.nsys = nsys;
    I.nb = lay.tree.nleaf;
    build_heat_levels(lay, hp, nthreads, I.host);
    I.nlev = (int)I.host.size();
    {
        // explicit diffusion limit 1 / max_i (2 kappa(r_i) sum_d 1/h_d^2), h = (dr, r dt, r sin(t) dp)
        const Tree &t = lay.tree;
        const int n = t.ni;
        double worst = 0.0;
        for (int bb = 0; bb < t.nleaf; bb++) {
            const Node &nd = t.nodes[t.leaf_node[bb]];
            const double dr = (nd.rng[0][1] - nd.rng[0][0]) / n, dth = (nd.rng[1][1] - nd.rng[1][0]) / n,
                         dph = (nd.rng[2][1] - nd.rng[2][0]) / n;
            for (int i = 1; i <= n; i++)
                for (int j = 1; j <= n; j++) {
                    const double r = (nd.rng[0][0] * (n - i + 0.5) + nd.rng[0][1] * (i - 0.5)) / n;
                    const double th = (nd.rng[1][0] * (n - j + 0.5) + nd.rng[1][1] * (j - 0.5)) / n;
                    const double h2 = r * dth, h3 = r * std::sin(th) * dph;
                    const double x = r / hp.r_ref;
                    const double kap = hp.kappa_power == 0 ? hp.kappa0 : hp.kappa0 / (x * x * x);
                    const double v = 2.0 * kap * (1.0 / (dr * dr) + 1.0 / (h2 * h2) + 1.0 / (h3 * h3));
                    worst = v > worst ? v : worst;
                }
        }
        I.dt_expl = 1.0 / worst;
    }
    if (!I.stream) HCHECK(hipStreamCreateWithFlags(&I.stream, hipStreamNonBlocking));
    I.L.resize(I.nlev);
    for (int l = 0; l < I.nlev; l++) {
        const HeatLevel &H = I.host[l];
        HLev &D = I.L[l];
        D.m = H.m; D.nb = I.nb;
        D.lm = 0;
        while ((1 << D.lm) < H.m) D.lm++;
        if (D.lm > 4) throw std::runtime_error("heat blocks are at most 16^3");
        D.M3 = (long long)H.m * H.m * H.m;
        if ((long long)I.nb * D.M3 > 0x7fffffffLL) throw std::runtime_error("heat level exceeds 2^31 cells");
        D.V = I.upload(H.V); D.Ai = I.upload(H.Ai); D.Aj = I.upload(H.Aj); D.Ak = I.upload(H.Ak);
        D.off = I.upload(H.offsum);
        std::vector<int2> gsw(H.g_src.size());
        for (size_t g = 0; g < gsw.size(); g++) gsw[g] = make_int2(H.g_src[g], H.g_wid[g]);
        D.gsw = I.upload(gsw);
        D.gw = I.upload(H.g_w);
    }
    I.Z.resize(I.nlev); I.Z2.resize(I.nlev); I.R.resize(I.nlev);
    for (int l = 0; l < I.nlev; l++) {
        I.Z[l] = I.dalloc(I.ist(l) * nsys);
        I.Z2[l] = I.dalloc(I.ist(l) * nsys);
        I.R[l] = l ? I.dalloc(I.ist(l) * nsys) : nullptr;
    }
    {
        // the coarsest level's face ghosts are copies of the donors' single cells
        const HeatLevel &H = I.host[I.nlev - 1];
        if (H.m != 1) throw std::runtime_error("the coarsest heat level must have one cell per block");
        for (int32_t w : H.g_wid)
            if (w >= 0) throw std::runtime_error("weighted face ghost on the 1-cell level");
    }
    I.coarse_work = I.dalloc((size_t)nsys * 3 * I.nb);
    I.x = I.dalloc(I.ist(0) * nsys);
    I.r = I.dalloc(I.ist(0) * nsys);
    I.q = I.dalloc(I.ist(0) * nsys);
    I.b = I.dalloc(I.ist(0) * nsys);
    I.tmp = I.dalloc(I.ist(0) * nsys);
    I.p = I.dalloc(I.ist(0) * nsys);
    I.inv_dt = I.dalloc(nsys);
    I.partial = I.dalloc((size_t)nsys * 3 * std::max(I.march_wgs(0), I.elem_wgs()));
    I.dots = I.dalloc((size_t)nsys * 3);
    I.sc = I.dalloc((size_t)nsys * kSc);
    void *a = nullptr;
    HCHECK(hipMalloc(&a, sizeof(int) * nsys));
    I.allocs.push_back(a);
    I.active = (int *)a;
    I.set_all_active();
    HCHECK(hipMalloc(&a, sizeof(int) * (nsys + 1)));
    I.allocs.push_back(a);
    I.iters_dev = (int *)a;
    I.nactive_dev = I.iters_dev + nsys;
    I.relres_dev = I.dalloc(nsys);
    I.drop_graphs();
    I.dt.assign(nsys, 1.0);
}

int HeatSolver::levels() const { return d->nlev; }
int HeatSolver::nsys() const { return d->nsys; }
long long HeatSolver::cells() const { return (long long)d->nb * d->L[0].M3; }
int HeatSolver::m() const { return d->L[0].m; }
hipStream_t HeatSolver::stream() const { return d->stream; }
void HeatSolver::set_corrector(HeatCorrector *c) {
    d->corr = c;
    d->drop_graphs();
}
void HeatSolver::set_graph(bool on) { d->use_graph = on; }
void HeatSolver::set_learned_blockstep(bool on) {
    d->blockstep = on;
    d->drop_graphs();
}
void HeatSolver::set_learned_iters(int n) {
    if (n < 0) throw std::runtime_error("the number of learned directions must be >= 0");
    d->learned_iters = n;
}
void HeatSolver::set_coarse(double tol, int max_iter) {
    if (!(tol >= 0.0) || max_iter < 0) throw std::runtime_error("coarse CG: tol >= 0 and max_iter >= 0");
    d->coarse_tol = tol;
    d->coarse_max = max_iter;
    d->drop_graphs();
}
void HeatSolver::set_learned_sweeps(int pre0, int post0) {
    if (pre0 < 0 || post0 < 0 || pre0 > 8 || post0 > 8) throw std::runtime_error("sweeps must be in [0, 8]");
    d->learned_pre0 = pre0;
    d->learned_post0 = post0;
    d->drop_graphs();
}

double HeatSolver::explicit_dt() const { return d->dt_expl; }

void HeatSolver::set_dt(const std::vector<double> &dt) {
    Impl &I = *d;
    if ((int)dt.size() != I.nsys) throw std::runtime_error("one dt per system");
    I.dt = dt;
    std::vector<double> inv(I.nsys);
    for (int s = 0; s < I.nsys; s++) inv[s] = 1.0 / dt[s];
    HCHECK(hipMemcpy(I.inv_dt, inv.data(), sizeof(double) * I.nsys, hipMemcpyHostToDevice));
}

void HeatSolver::set_rhs_from_T(const std::vector<double> &T) {
    Impl &I = *d;
    if ((long long)T.size() != I.ist(0) * I.nsys) throw std::runtime_error("T has the wrong size");
    HCHECK(hipMemcpy(I.tmp, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_hrhs, I.grid_cells(0), dim3(256), 0, I.stream, I.L[0], (const double *)I.tmp, I.b,
                       (const double *)I.inv_dt, I.ist(0));
    HCHECK(hipStreamSynchronize(I.stream));
}

void HeatSolver::set_rhs(const std::vector<double> &b) {
    Impl &I = *d;
    if ((long long)b.size() != I.ist(0) * I.nsys) throw std::runtime_error("b has the wrong size");
    HCHECK(hipMemcpy(I.b, b.data(), sizeof(double) * b.size(), hipMemcpyHostToDevice));
}

void HeatSolver::pfss_field(const std::vector<double> &bot, const std::vector<double> &geo, std::vector<double> &B) {
    Impl &I = *d;
    const long long n = I.ist(0);
    if ((long long)bot.size() != (long long)I.nb * I.L[0].m * I.L[0].m ||
        (long long)geo.size() != (long long)I.nb * (2 * I.L[0].m + 3))
        throw std::runtime_error("pfss_field: table sizes do not match the layout");
    double *dbot = nullptr, *dgeo = nullptr, *dB = nullptr;
    HCHECK(hipMalloc(&dbot, sizeof(double) * bot.size()));
    HCHECK(hipMalloc(&dgeo, sizeof(double) * geo.size()));
    HCHECK(hipMalloc(&dB, sizeof(double) * 3 * n));
    HCHECK(hipMemcpy(dbot, bot.data(), sizeof(double) * bot.size(), hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(dgeo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pfss_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, I.stream, I.L[0],
                       (const double *)I.x, (const double *)dbot, (const double *)dgeo, dB);
    HCHECK(hipGetLastError());
    B.resize(3 * n);
    HCHECK(hipStreamSynchronize(I.stream));
    HCHECK(hipMemcpy(B.data(), dB, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    HCHECK(hipFree(dbot));
    HCHECK(hipFree(dgeo));
    HCHECK(hipFree(dB));
}

void HeatSolver::apply(const std::vector<double> &x, std::vector<double> &y) {
    // y = A x for system-major interior vectors (parity tests)
    Impl &I = *d;
    const long long n = I.ist(0) * I.nsys;
    if ((long long)x.size() != n) throw std::runtime_error("x has the wrong size");
    I.set_all_active();
    HCHECK(hipMemcpy(I.p, x.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    I.apply0(I.p, I.q, nullptr);
    y.resize(n);
    HCHECK(hipStreamSynchronize(I.stream));
    HCHECK(hipMemcpy(y.data(), I.q, sizeof(double) * n, hipMemcpyDeviceToHost));
}

void HeatSolver::vcycle_host(const std::vector<double> &r, std::vector<double> &z) {
    Impl &I = *d;
    const long long n = I.ist(0) * I.nsys;
    if ((long long)r.size() != n) throw std::runtime_error("r has the wrong size");
    I.set_all_active();
    HCHECK(hipMemcpy(I.q, r.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    double *zp = nullptr;
    I.vcycle(I.q, zp, I.pre, I.post, nullptr);
    z.resize(n);
    HCHECK(hipStreamSynchronize(I.stream));
    HCHECK(hipMemcpy(z.data(), zp, sizeof(double) * n, hipMemcpyDeviceToHost));
}

HeatResult HeatSolver::solve(double tol, int max_iter, bool use_corrector) {
    return solve(tol, max_iter, use_corrector ? kSolveLearned : kSolvePlain, nullptr);
}

void HeatSolver::Impl::ensure_learned() {
    if (rp) return;
    rp = dalloc(ist(0) * nsys);
    cc = dalloc(ist(0) * nsys);
    ablk = dalloc((size_t)nsys * nb);
    void *a = nullptr;
    HCHECK(hipMalloc(&a, sizeof(float) * ist(0) * nsys));
    allocs.push_back(a);
    xin = (float *)a;
}

// r' = r - A z, network input r'/diag and its per-system scale (root mean
// square) in sc[3]
void HeatSolver::Impl::features(double *z) {
    launch_march(0, z, OpFeat{rp, xin, r, nullptr}, partial);
    final_dots(1, march_wgs(0));
    scalars(3, (double)ist(0));
}

// c = scale * net(features / scale) and its step a = <c,r'>/<c,Ac> (sc[4]);
// unless first, beta = <z + a c, q>/gamma of the direction update (sc[1])
void HeatSolver::Impl::learned_correction(bool first) {
    corr->infer_heat(xin, sc + 3, kSc, nb, ist(0), cc, (long long)nsys * nb, active, stream);
    launch_march(0, cc, OpApplyDot{rp, q}, partial);
    if (blockstep) {
        const unsigned nwg = march_wgs(0);
        hipLaunchKernelGGL(k_hblockstep, dim3((nwg + 255) / 256, (unsigned)nsys), dim3(256), 0, stream, (int)nwg, nb,
                           partial, ablk, (const int *)active);
        if (!first) {
            final_dots(1, nwg);
            scalars(7);
        }
        return;
    }
    final_dots(3, march_wgs(0));
    scalars(2);
    if (!first) scalars(5);
}

HeatResult HeatSolver::solve(double tol, int max_iter, int mode, HeatRecord *rec) {
    Impl &I = *d;
    const int ns = I.nsys;
    const long long ni = I.ist(0);
    const bool learned = mode == kSolveLearned, recording = mode == kSolveRecord;
    if (learned && !I.corr) throw std::runtime_error("no learned correction is attached to the heat solver");
    if ((learned || recording) && I.L[0].m != 16) throw std::runtime_error("the learned correction needs 16^3 blocks");
    if (recording && !rec) throw std::runtime_error("record mode needs a record buffer");
    if (learned || recording) I.ensure_learned();
    HeatResult res;
    res.iters.assign(ns, 0);
    res.relres.assign(ns, 0.0);
    I.set_all_active();
    HCHECK(hipStreamSynchronize(I.stream));
    const auto t0 = std::chrono::steady_clock::now();
    const dim3 g0 = I.grid_cells(0);
    const long long rec0 = rec ? rec->count : 0;
    // x = 0, r = b, <b, b>
    I.zero(I.x, ni * ns);
    hipLaunchKernelGGL(k_hcopy, dim3((unsigned)((ni * ns + 255) / 256)), dim3(256), 0, I.stream, (const double *)I.b, I.r,
                       ni * ns);
    hipLaunchKernelGGL(k_hdot, I.grid_elem(), dim3(256), 0, I.stream, ni, (const double *)I.b, (const double *)I.b,
                       (const double *)nullptr, (const double *)nullptr, 1, (const int *)nullptr, I.partial, ni);
    I.final_dots(1, I.elem_wgs());
    I.scalars(4);
    std::vector<double> bn(3 * ns);
    HCHECK(hipMemcpyAsync(bn.data(), I.dots, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, I.stream));
    HCHECK(hipStreamSynchronize(I.stream));
    // a system with b = 0 has the solution 0: it never becomes active
    std::vector<int> act(ns, 1);
    int nact = 0;
    for (int s = 0; s < ns; s++) {
        if (bn[3 * s] == 0.0) act[s] = 0;
        nact += act[s];
    }
    HCHECK(hipMemcpy(I.active, act.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_hizero, dim3((ns + 255) / 256), dim3(256), 0, I.stream, I.iters_dev, ns);

    auto record = [&](const double *z) {
        // one snapshot (residual/error pair) per system still iterating, each
        // contributing per_snapshot random blocks, while room is left
        HCHECK(hipMemcpyAsync(act.data(), I.active, sizeof(int) * ns, hipMemcpyDeviceToHost, I.stream));
        HCHECK(hipStreamSynchronize(I.stream));
        std::vector<int> rs, rb;
        for (int s = 0; s < ns; s++) {
            if (!act[s]) continue;
            if (rec->count + (long long)rs.size() + rec->per_snapshot > rec->capacity) break;
            for (int k = 0; k < rec->per_snapshot; k++) {
                rec->rng = rec->rng * 6364136223846793005ULL + 1442695040888963407ULL;
                rs.push_back(s);
                rb.push_back((int)((rec->rng >> 11) % (uint64_t)I.nb));
            }
            rec->pairs++;
        }
        const int R = (int)rs.size();
        if (R == 0) return;
        HCHECK(hipMemcpyAsync(rec->rs + rec->count, rs.data(), sizeof(int) * R, hipMemcpyHostToDevice, I.stream));
        HCHECK(hipMemcpyAsync(rec->rb + rec->count, rb.data(), sizeof(int) * R, hipMemcpyHostToDevice, I.stream));
        hipLaunchKernelGGL(k_hrecord, dim3(R), dim3(256), 0, I.stream, I.L[0], (const float *)I.xin, (const double *)I.x, z,
                           (const int *)(rec->rs + rec->count), (const int *)(rec->rb + rec->count), rec->X + rec->count * 4096,
                           rec->W + rec->count * 4096, (const double *)I.sc, rec->scale + rec->count, ni);
        HCHECK(hipStreamSynchronize(I.stream));   // rs/rb host vectors leave scope
        rec->count += R;
    };
    // preconditioner and direction update; first: the initial direction p = z;
    // lrn: learned correction on the V-cycle (else the plain V-cycle).  Both
    // orthogonalise against the previous direction with <z', q>/<p, q>, so a
    // solve can switch from one to the other between iterations.
    auto precond_direction = [&](bool first, bool lrn) {
        double *z = nullptr;
        if (lrn) {
            const unsigned nwg = I.vcycle(I.r, z, I.learned_pre0, I.learned_post0, first ? nullptr : I.q);
            if (!first) {
                I.final_dots(1, nwg);
                I.scalars(6);                // <z, q>
            }
            I.features(z);
            I.learned_correction(first);     // step a and (unless first) beta
            hipLaunchKernelGGL(k_hupdate_p2, g0, dim3(256), 0, I.stream, ni, I.p, (const double *)z, (const double *)I.cc,
                               (const double *)I.sc, (const double *)(I.blockstep ? I.ablk : nullptr), I.nb,
                               (const int *)I.active, ni, first ? 0 : 1);
            return;
        }
        const int pre0 = recording ? I.learned_pre0 : I.pre, post0 = recording ? I.learned_post0 : I.post;
        const unsigned nwg = I.vcycle(I.r, z, pre0, post0, first ? nullptr : I.q);
        if (!first) {
            I.final_dots(1, nwg);
            I.scalars(1);
        }
        if (recording) {
            I.features(z);
            record(z);
        }
        hipLaunchKernelGGL(k_hupdate_p, g0, dim3(256), 0, I.stream, ni, I.p, (const double *)z, (const double *)I.sc,
                           (const int *)I.active, ni, first ? 0 : 1);
    };
    // one FCG iteration; the convergence test runs on the device (k_hconverge)
    auto body = [&](bool lrn) {
        I.apply0(I.p, I.q, I.r);                         // q = A p with <p,q>, <p,r>
        I.final_dots(2, I.march_wgs(0));
        I.scalars(0);
        hipLaunchKernelGGL(k_hupdate_xr, I.grid_elem(), dim3(256), 0, I.stream, ni, I.x, I.r, (const double *)I.p,
                           (const double *)I.q, (const double *)I.sc, (const int *)I.active, I.partial, ni);
        I.final_dots(1, I.elem_wgs());
        hipLaunchKernelGGL(k_hconverge, dim3(1), dim3(256), 0, I.stream, ns, (const double *)I.dots, (const double *)I.sc,
                           tol, I.active, I.iters_dev, I.relres_dev, I.nactive_dev);
        precond_direction(false, lrn);
    };
    // directions 0 .. nlearn-1 come from the learned correction, later ones
    // from the plain V-cycle
    const int nlearn = learned ? (I.learned_iters > 0 ? I.learned_iters : INT_MAX) : 0;
    if (nact > 0) {
        if (learned) I.corr->prepare(I.stream);
        precond_direction(true, nlearn > 0);
    }
    const bool graph = I.use_graph && !recording;
    // one FCG iteration per kind of direction captured as a hipGraph and
    // replayed: the host reads one counter per iteration.  Arithmetic equals
    // the eager launches'.
    auto graph_of = [&](bool lrn) -> hipGraphExec_t {
        const int gi = lrn ? 1 : 0;
        hipGraphExec_t &G = I.graph[gi];
        if (G && I.graph_tol[gi] != tol) {
            HCHECK(hipGraphExecDestroy(G));
            G = nullptr;
        }
        if (!G) {
            hipGraph_t g;
            HCHECK(hipStreamBeginCapture(I.stream, hipStreamCaptureModeThreadLocal));
            body(lrn);
            HCHECK(hipStreamEndCapture(I.stream, &g));
            HCHECK(hipGraphInstantiate(&G, g, nullptr, nullptr, 0));
            HCHECK(hipGraphDestroy(g));
            I.graph_tol[gi] = tol;
        }
        return G;
    };
    int launched = 0;
    while (nact > 0 && launched < max_iter) {
        const bool lrn = launched + 1 < nlearn;          // kind of the direction this iteration builds
        if (graph) HCHECK(hipGraphLaunch(graph_of(lrn), I.stream));
        else body(lrn);
        launched++;
        HCHECK(hipMemcpyAsync(&nact, I.nactive_dev, sizeof(int), hipMemcpyDeviceToHost, I.stream));
        HCHECK(hipStreamSynchronize(I.stream));
    }
    std::vector<int> itd(ns);
    std::vector<double> rel(ns);
    HCHECK(hipMemcpy(itd.data(), I.iters_dev, sizeof(int) * ns, hipMemcpyDeviceToHost));
    HCHECK(hipMemcpy(rel.data(), I.relres_dev, sizeof(double) * ns, hipMemcpyDeviceToHost));
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int s = 0; s < ns; s++) {
        const bool zero_rhs = bn[3 * s] == 0.0;
        res.iters[s] = zero_rhs ? 0 : itd[s];
        res.relres[s] = zero_rhs ? 0.0 : rel[s];
    }
    if (rec && rec->count > rec0) {
        // targets from the converged solution: the error each recorded V-cycle left
        const long long nrec = rec->count - rec0;
        hipLaunchKernelGGL(k_htarget, dim3((unsigned)nrec), dim3(256), 0, I.stream, I.L[0], (const double *)I.x,
                           (const int *)(rec->rs + rec0), (const int *)(rec->rb + rec0), (const double *)(rec->scale + rec0),
                           (const double *)(rec->W + rec0 * 4096), rec->Y + rec0 * 4096, ni);
        HCHECK(hipStreamSynchronize(I.stream));
    }
    HCHECK(hipGetLastError());
    return res;
}

void HeatSolver::get_solution(int s, std::vector<double> &T) const {
    const Impl &I = *d;
    T.resize(I.ist(0));
    HCHECK(hipStreamSynchronize(I.stream));
    HCHECK(hipMemcpy(T.data(), I.x + s * I.ist(0), sizeof(double) * I.ist(0), hipMemcpyDeviceToHost));
}

void HeatSolver::get_rhs(int s, std::vector<double> &b) const {
    const Impl &I = *d;
    b.resize(I.ist(0));
    HCHECK(hipMemcpy(b.data(), I.b + s * I.ist(0), sizeof(double) * I.ist(0), hipMemcpyDeviceToHost));
}

}  // namespace falcon
