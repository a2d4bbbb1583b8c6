// Device kernels of the RK4 hot path for gfx950.
//
// Layout in HBM: every state buffer is [block][var][I][J][K] with the padded
// block (NI,NJ,NK) = (ni+2ng, nj+2ng, nk+2ng) contiguous and K (phi) fastest,
// so a wavefront's loads along phi are contiguous.
//
// Stage and diffusion kernels march each block along r.  A workgroup of 256
// lanes owns a 16x16 (theta, phi) tile of one block; lane (ty, tx) owns one
// (j, k) column and keeps a 5-plane register window of all 8 variables for the
// r-stencils, while the current r-plane with its 2-cell theta/phi halo is staged
// in double-buffered LDS for the theta/phi stencils.  Corner (edge) halo cells
// are never needed by these stencils, so they are never loaded.
//
// Every arithmetic expression keeps the operand order of the reference line it
// restates; with -ffp-contract=off the FP64 instantiation is bit-identical to
// the reference.  The FP32 instantiation (T = float) is the deviation path: the
// stencil arithmetic runs in FP32 while the stage accumulation into the FP64
// state, the ghost interpolation weights and the time step stay FP64.
#include <algorithm>
#include <stdexcept>

#include "device.hpp"

namespace falcon {

// Stage, diffusion and dt kernels take the (theta, phi) tile edge TL as a
// template parameter: 16 (256 lanes, one 16x16 tile per 16^3 block face) or 8
// (64 lanes).  Each cell's arithmetic is the same for either edge, so the
// choice (made by timing both, falcon_autotune) never changes a result.
template <int TL>
constexpr int tile_pl() { return TL + 4; }   // LDS plane edge (tile + 2 halo each side)

// Tile of workgroup w of a one-workgroup-per-tile launch of n tiles.  The
// dispatcher deals workgroups to the 8 XCDs in turn (w mod 8), and each XCD
// has its own L2.  Giving XCD x the contiguous tile range [x*n/8, ...) in
// dispatch order puts phi-adjacent blocks (consecutive ids) on the same XCD at
// the same time, so the halo each reads from the other is an L2 hit: on the
// 50 M-cell shell the FP32 stage reads 68 instead of 134 bytes per cell from
// HBM (stages 1-3) and the FP32 diffusion 158 instead of 220.
__device__ __forceinline__ int xcd_tile(int w, int n) {
    const int q = n >> 3, r = n & 7, x = w & 7, k = w >> 3;
    return x * q + min(x, r) + k;
}

// the lane index of a flat 1-D launch with the workgroups in xcd_tile's order:
// a run of consecutive lanes (the cells of one block) is served by one XCD,
// whose L2 then holds the neighbour planes the run shares.  The scattered
// gathers of the ghost fill measure 5-8% slower in this order than in the
// dispatcher's round-robin, and keep the latter.
__device__ __forceinline__ long long xcd_lane() {
    return (long long)xcd_tile((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
}

// psi (the ninth variable of the reference's MHD blocks, 3_6/ModBlock.f90:163)
// is one padded variable per block, [block][P] in the cell_off layout; the
// ghost tables address [block][8][P] state buffers, so an entry's offset o
// maps to (o / 8P) P + o mod 8P.
__device__ __forceinline__ long long psi_off(const Dims &d, long long o) {
    const long long P8 = 8 * d.P, b = o / P8;
    return b * d.P + (o - b * P8);
}

// radial coefficient rows the stage kernel stages in LDS, and the largest padded
// radial extent NI a march accepts (the launchers reject larger blocks)
constexpr int kStageRows = 10;

constexpr int kMaxNI = 132;

// ModDerivative_1st_O4_3D (3_6/ModDerivative.f90:16-19):
// (1./12./dx)*( -f(+2) + 8.*f(+1) - 8.*f(-1) + f(-2) )
template <typename T>
__device__ __forceinline__ T d4(T fm2, T fm1, T fp1, T fp2, T c) {
    return c * ((((-fp2) + T(8) * fp1) - T(8) * fm1) + fm2);
}

// ModLinReconstruct_minmod (3_6/ModLinReconstruct.f90:37-48) for one triple
template <typename T>
__device__ __forceinline__ T minmod(T ul, T u, T ur) {
    T a = (ur - ul) * T(0.5);
    T b = T(2) * (ur - u);
    T c = T(2) * (u - ul);
    T ab = a * b, bc = b * c;
    T m1 = ab > T(0) ? ab : T(0);
    T m2 = bc > T(0) ? bc : T(0);
    T s = (m1 * m2) * T(1.0e10) - T(1.0e-30f);
    T half = copysign(T(0.5), s);
    T fa = fabs(a), fb = fabs(b), fc = fabs(c);
    T mn = fa < fb ? fa : fb;
    mn = mn < fc ? mn : fc;
    return (half + T(0.5)) * copysign(mn, a);
}

// One face of ModDiffusion_Aritificial_1 (3_6/ModDiffusion.f90:50-59) between
// cell y (value b) and cell x = y+1 (value c); a = y-1, d = x+1.
// MAX(0.0, NaN) is taken as 0.0 (the comparison below is false for NaN).
template <typename T>
__device__ __forceinline__ T face_flux(T a, T b, T c, T d) {
    T dX = minmod(b, c, d);
    T dY = minmod(a, b, c);
    T num = ((c - dX * T(0.5)) - b) - dY * T(0.5);
    T ph = T(1) + T(2) * (num / (c - b) - T(1));
    ph = ph > T(0) ? ph : T(0);
    return ((-T(0.5)) * ph) * num;
}

// The same face from the two cells' minmod slopes: sY at y (value b), sX at
// x (value c).  The reference forms every cell's slope once per direction
// (ModLinReconstruct_minmod over the whole array) and both faces of a cell use
// it; face_flux recomputes it per face.  RCP (FP32 path): the limiter ratio as
// a product with the hardware reciprocal of (c - b), whose 1 ulp error is far
// below the FP32 storage rounding; 0 and +-0 denominators give the same
// infinities and NaN as the quotient, so the NaN -> 0 rule still applies.
template <typename T, bool RCP>
__device__ __forceinline__ T face_flux_s(T b, T c, T sY, T sX) {
    T num = ((c - sX * T(0.5)) - b) - sY * T(0.5);
    T q = RCP ? num * (T)__builtin_amdgcn_rcpf((float)(c - b)) : num / (c - b);
    T ph = T(1) + T(2) * (q - T(1));
    ph = ph > T(0) ? ph : T(0);
    return ((-T(0.5)) * ph) * num;
}

// Shared r-marching skeleton: loads the column window and the halo'd plane.
// With a pull table (the row of block b: entry (di+1)*9+(dj+1)*3+(dk+1) is the
// same-level neighbour whose interior holds this block's ghosts in direction
// (di,dj,dk), or -1), those ghosts are read from the neighbour at the ghost
// index minus (di*ni, dj*nj, dk*nk); the radial ghost planes of the column are
// direction (-1,0,0) / (+1,0,0), the theta/phi halo strips (0,dj,0) / (0,0,dk).
// PULL = false compiles the plain reads (the pull table is ignored).
template <typename T, typename TIN, bool PULL, int TL>
struct Marcher {
    const TIN *blk;
    long long dlo, dhi;     // offset from blk of the column source of the radial ghost
                            // planes below / above the block, plane shift included
    const TIN *hblk;        // halo strip source
    long long P;
    int SI, SJ;             // main-part plane and row pitch (cell_off)
    int ilo, ihi;           // first interior plane, first plane above the interior
    int J, K;           // padded column of this lane
    long long coff;     // J*SJ + K, clamped into the block for lanes outside a partial tile
    bool act;
    bool hv;
    long long hoff;     // cell_off of the halo point at plane 0 (a valid cell for lanes without one)
    int hsi;            // plane pitch of the halo point's part (main or side)
    int hly, hlx;

    __device__ __forceinline__ void init(const Dims &d, const TIN *in, int b, int j0, int k0, int TJa, int TKa,
                         const int32_t *pull = nullptr) {
        P = d.P;
        SI = (int)d.SI;
        SJ = d.SJ;
        blk = in + (long long)b * 8 * P;
        ilo = d.ng;
        ihi = d.ng + d.ni;
        hblk = blk;
        dlo = dhi = 0;
        if (PULL && pull) {
            const int Dlo = pull[b * 27 + 4], Dhi = pull[b * 27 + 22];
            if (Dlo >= 0) dlo = (long long)(Dlo - b) * 8 * P + (long long)d.ni * SI;
            if (Dhi >= 0) dhi = (long long)(Dhi - b) * 8 * P - (long long)d.ni * SI;
        }
        const int tx = threadIdx.x & (TL - 1), ty = threadIdx.x / TL;
        act = ty < TJa && tx < TKa;
        J = j0 + ty + d.ng;
        K = k0 + tx + d.ng;
        // loads of inactive lanes read the tile's first cell, so every load can be
        // issued without a branch; their values are never used
        coff = act ? (long long)J * SJ + (K - d.ng) : (long long)(j0 + d.ng) * SJ + k0;
        const int h = threadIdx.x;
        hv = false; hly = 0; hlx = 0;
        if (h < 4 * TL) {
            int r = h / TL, kk = h % TL;
            hly = r < 2 ? r : TJa + r;
            hlx = kk + 2;
            hv = kk < TKa;
        } else if (h < 8 * TL) {
            int h2 = h - 4 * TL;
            int c = h2 / TL, jj = h2 % TL;
            hlx = c < 2 ? c : TKa + c;
            hly = jj + 2;
            hv = jj < TJa;
        }
        const int HJ = j0 + hly, HK = k0 + hlx;   // padded halo cell
        const bool hmain = HK >= d.ng && HK < d.ng + d.nk;
        hoff = hv ? cell_off(d, 0, HJ, HK) : coff;
        hsi = (!hv || hmain) ? SI : d.NJ * 2 * d.ng;
        if (PULL && pull && hv) {
            const int dj = HJ < d.ng ? -1 : (HJ >= d.ng + d.nj ? 1 : 0);
            const int dk = HK < d.ng ? -1 : (HK >= d.ng + d.nk ? 1 : 0);
            if (dj | dk) {
                const int D = pull[b * 27 + 9 + (dj + 1) * 3 + (dk + 1)];
                if (D >= 0) {
                    // the neighbour's interior cell: always in its main part
                    hblk = in + (long long)D * 8 * P;
                    hoff = cell_off(d, 0, HJ - dj * d.nj, HK - dk * d.nk);
                    hsi = SI;
                }
            }
        }
    }
    // I is the same for every lane of the workgroup, 0 <= I < NI
    __device__ __forceinline__ T col(int v, int I) const {
        long long o = (long long)v * P + (long long)I * SI + coff;
        if (PULL) o += (I < ilo ? dlo : 0ll) + (I >= ihi ? dhi : 0ll);
        return (T)blk[o];
    }
    // interior planes only
    __device__ __forceinline__ T halo(int v, int I) const {
        return (T)(PULL ? hblk : blk)[(long long)v * P + (long long)I * hsi + hoff];
    }
};

// ---------------------------------------------------------------- stage

// RHS of ModEquation_Dynamo (3_6/ModEquation.f90:24-162) for one cell, in
// the reference's operation order, computed in T (FP64 on both paths).
// W(v, q): variable v at r-offset q-2 of this lane's column, read from the
// register window `win` of storage type TW (the FP32 path keeps its stage
// values in FP32 and widens them here); L(v,dj,dk): the LDS plane, widened the
// same way.
// RECIP (FP32 path only): the divisions by rho0 and by sin(theta) become
// products with irho0 = 1/rho0 and sinv = 1/sin(theta), the latter applied
// after the phi derivative instead of to each of its four points.  Each is a
// rounding-level change in FP64 (relative 1e-16, far below the FP32 storage
// rounding of the stage values) and removes 14 FP64 divisions per cell.
template <typename T, typename TW, bool RECIP, typename LDSF>
__device__ __forceinline__ void cell_rhs(const TW (&win)[8][5], LDSF L, const T xr[5], const T rho0r[5], const T gpr[5],
                                         const T gqr[5], const T st[5], T rinv, T r2inv, T xir2, T heat, T gor,
                                         T sinv, T cot, T cr, T ct, T cp, T irho0, T R[8]) {
    auto W = [&](int v, int q) -> T { return (T)win[v][q]; };
    const T rho0c = rho0r[2];
    auto over_rho0 = [&](T x) -> T { return RECIP ? x * irho0 : x / rho0c; };
    // d4 of f/sin(theta) along phi (sin(theta) is constant along phi)
    auto d4_over_st = [&](T f0, T f1, T f3, T f4) -> T {
        return RECIP ? d4(f0, f1, f3, f4, cp) * sinv : d4(f0 / st[2], f1 / st[2], f3 / st[2], f4 / st[2], cp);
    };
    const T vr = W(1, 2), vt = W(2, 2), vp = W(3, 2);
    const T br = W(4, 2), bt = W(5, 2), bp = W(6, 2);

    // Mass conservation (:50-61) with ModSpherical_div (:104-159)
    {
        T w[5];
#pragma unroll
        for (int q = 0; q < 5; q++) w[q] = (rho0r[q] * W(1, q)) * (xr[q] * xr[q]);
        T t1 = d4(w[0], w[1], w[3], w[4], cr) * r2inv;
        T t2 = d4((rho0c * L(2, -2, 0)) * st[0], (rho0c * L(2, -1, 0)) * st[1], (rho0c * L(2, 1, 0)) * st[3],
                  (rho0c * L(2, 2, 0)) * st[4], ct);
        T dd = d4(rho0c * L(3, 0, -2), rho0c * L(3, 0, -1), rho0c * L(3, 0, 1), rho0c * L(3, 0, 2), cp);
        t2 = t2 + dd;
        t2 = t2 * sinv;
        t2 = t2 * rinv;
        R[0] = xir2 * (t1 + t2);
    }
    // A_dot_Grad_f (ModSpherical.f90:241-279) with A = v at the centre
    auto adg = [&](int f) -> T {
        T o = vr * d4(W(f, 0), W(f, 1), W(f, 3), W(f, 4), cr);
        T tmp = vt * d4(L(f, -2, 0), L(f, -1, 0), L(f, 1, 0), L(f, 2, 0), ct);
        o = o + tmp * rinv;
        tmp = (vp * d4(L(f, 0, -2), L(f, 0, -1), L(f, 0, 1), L(f, 0, 2), cp)) * rinv;
        o = o + sinv * tmp;
        return o;
    };
    // Inertial force (:66-74): -A_dot_nabla_B(v,v) (ModSpherical.f90:281-343)
    {
        T a1 = adg(1), a2 = adg(2), a3 = adg(3);
        T tr = (-vt) * vt - vp * vp;
        a1 = a1 + tr * rinv;
        T tt = (-vp) * vp;
        tt = tt * cot;
        tt = tt + vt * vr;
        a2 = a2 + tt * rinv;
        T tp = vp * vt;
        tp = tp * cot;
        tp = tp + vp * vr;
        a3 = a3 + tp * rinv;
        R[1] = -a1; R[2] = -a2; R[3] = -a3;
    }
    // Pressure gradient (:79-92): p1 (:40-46), ModSpherical_Grad_f (:201-239)
    {
        T pr[5];
#pragma unroll
        for (int q = 0; q < 5; q++) pr[q] = gpr[q] * W(0, q) + gqr[q] * W(7, q);
        auto p1L = [&](int dj, int dk) -> T { return gpr[2] * L(0, dj, dk) + gqr[2] * L(7, dj, dk); };
        T t1 = d4(pr[0], pr[1], pr[3], pr[4], cr);
        T t2 = d4(p1L(-2, 0), p1L(-1, 0), p1L(1, 0), p1L(2, 0), ct);
        T t3 = d4(p1L(0, -2), p1L(0, -1), p1L(0, 1), p1L(0, 2), cp);
        T g2 = t2 * rinv;
        T g3 = (t3 * rinv) * sinv;
        R[1] = R[1] - over_rho0(t1);
        R[2] = R[2] - over_rho0(g2);
        R[3] = R[3] - over_rho0(g3);
    }
    // Lorentz force (:97-116): curl(B) (ModSpherical.f90:38-102) x B / rho0
    {
        T o1 = d4(L(6, -2, 0) * st[0], L(6, -1, 0) * st[1], L(6, 1, 0) * st[3], L(6, 2, 0) * st[4], ct) -
               d4(L(5, 0, -2), L(5, 0, -1), L(5, 0, 1), L(5, 0, 2), cp);
        T o2 = d4_over_st(L(4, 0, -2), L(4, 0, -1), L(4, 0, 1), L(4, 0, 2)) -
               d4(W(6, 0) * xr[0], W(6, 1) * xr[1], W(6, 3) * xr[3], W(6, 4) * xr[4], cr);
        T o3 = d4(W(5, 0) * xr[0], W(5, 1) * xr[1], W(5, 3) * xr[3], W(5, 4) * xr[4], cr) -
               d4(L(4, -2, 0), L(4, -1, 0), L(4, 1, 0), L(4, 2, 0), ct);
        o1 = o1 * rinv; o2 = o2 * rinv; o3 = o3 * rinv;
        o1 = o1 * sinv;
        T c1 = o2 * bp - o3 * bt;
        T c2 = o3 * br - o1 * bp;
        T c3 = o1 * bt - o2 * br;
        R[1] = R[1] + over_rho0(c1);
        R[2] = R[2] + over_rho0(c2);
        R[3] = R[3] + over_rho0(c3);
    }
    // Gravity (:121-127)
    R[1] = R[1] - gor * W(0, 2);
    // Entropy advection and heating (:132-151)
    R[7] = -adg(7);
    R[7] = R[7] + heat;
    // Induction (:154-162): curl(v x B), cross product per point (ModSpherical.f90:26-36)
    {
        auto e1L = [&](int dj, int dk) -> T { return L(2, dj, dk) * L(6, dj, dk) - L(3, dj, dk) * L(5, dj, dk); };
        auto e2L = [&](int dj, int dk) -> T { return L(3, dj, dk) * L(4, dj, dk) - L(1, dj, dk) * L(6, dj, dk); };
        auto e3L = [&](int dj, int dk) -> T { return L(1, dj, dk) * L(5, dj, dk) - L(2, dj, dk) * L(4, dj, dk); };
        T e2r[5], e3r[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            e2r[q] = W(3, q) * W(4, q) - W(1, q) * W(6, q);
            e3r[q] = W(1, q) * W(5, q) - W(2, q) * W(4, q);
        }
        T o1 = d4(e3L(-2, 0) * st[0], e3L(-1, 0) * st[1], e3L(1, 0) * st[3], e3L(2, 0) * st[4], ct) -
               d4(e2L(0, -2), e2L(0, -1), e2L(0, 1), e2L(0, 2), cp);
        T o2 = d4_over_st(e1L(0, -2), e1L(0, -1), e1L(0, 1), e1L(0, 2)) -
               d4(e3r[0] * xr[0], e3r[1] * xr[1], e3r[3] * xr[3], e3r[4] * xr[4], cr);
        T o3 = d4(e2r[0] * xr[0], e2r[1] * xr[1], e2r[3] * xr[3], e2r[4] * xr[4], cr) -
               d4(e1L(-2, 0), e1L(-1, 0), e1L(1, 0), e1L(2, 0), ct);
        o1 = o1 * rinv; o2 = o2 * rinv; o3 = o3 * rinv;
        o1 = o1 * sinv;
        R[4] = o1; R[5] = o2; R[6] = o3;
    }
}

// T is the storage type of the stage input (the LDS planes and the register
// window); the right-hand side is computed in FP64 for either.  Stage values
// stored in FP32 are the FP32 path's only rounding: FP32 arithmetic on the
// stencils and FP32 copies of the metric and background coefficients were
// each measured to put 1e-6 of error into the 100-step FP32-vs-FP64 state on
// the 50 M-cell shell, FP32 storage alone 5e-7.
// MODE 0: out = U + (dt*R)/div on the interior (3_6/ModAdvance.f90:87-93), the
//         sum formed in FP64 from the base U of type TB (the FP64 state, or for
//         the FP32 path's intermediate stages its FP32 copy: those stage values
//         are stored in FP32 anyway); with SHADOW the same value is also stored
//         rounded to FP32 in `shadow`.
// MODE 1: write R (interior layout [block][var][i][j][k], FP64) for parity tests.
// pull (null: none): ghost cells of `in` in the directions it lists are read
// from the same-level neighbour's interior (Marcher).
// GLM (FP64 only): grad(psi) from `glm` is subtracted from R(B) (GLM
// divergence cleaning); a separate instantiation, so the plain one is unchanged.
// The body is shared by two kernels, k_stage and k_stage_w2 (below).
template <typename T, typename TIN, typename TOUT, typename TB, int MODE, bool SHADOW, int TL, bool PERSIST,
          bool GLM = false>
__device__ __forceinline__ void stage_body(Dims d, DevTables t, const TIN *__restrict__ in,
                                               const TB *__restrict__ U, TOUT *__restrict__ out,
                                               float *__restrict__ shadow, const double *__restrict__ dtp,
                                               double div, const int32_t *__restrict__ pull,
                                               const double *__restrict__ glm) {
    constexpr int PL = tile_pl<TL>();
    __shared__ T lds[2][8][PL][PL];
    // LEAN (FP32 storage): the block's radial coefficient rows (xi, rho0, gp,
    // gq, rinv, r2inv, xir2, heat, g/rho0, and 1/rho0) are staged in LDS once
    // per tile, U of the plane is loaded at the top of the plane, and the next
    // plane's loads are issued unconditionally from clamped addresses.  The
    // plane loop then waits on no global load in the middle of its arithmetic
    // (a load issued after the prefetch would wait for the prefetch too), which
    // took the FP32 stage from 2.99 to 2.78 ms per launch on the 50 M-cell
    // shell.  The FP64 instantiation keeps the table loads: with the LDS rows
    // its register demand exceeds two wavefronts per SIMD.
    constexpr bool LEAN = sizeof(T) == 4;
    __shared__ double crow[kStageRows][kMaxNI];
    const int ntk = (d.nk + TL - 1) / TL, tpb = ((d.nj + TL - 1) / TL) * ntk;
    const int ntot = d.nblocks * tpb;
    // PERSIST: a grid smaller than the tile count (a persistent launch) takes
    // the next tile at a stride of the grid size; otherwise one pass, tile =
    // the workgroup
    for (int tile = PERSIST ? (int)blockIdx.x : xcd_tile(blockIdx.x, ntot); tile < ntot; tile += gridDim.x) {
        const int b = tile / tpb;
        const int trem = tile - b * tpb;
        const int tj = trem / ntk, tk = trem - tj * ntk;
        const int j0 = tj * TL, k0 = tk * TL;
        const int TJa = min(TL, d.nj - j0), TKa = min(TL, d.nk - k0);
        const int tx = threadIdx.x & (TL - 1), ty = threadIdx.x / TL;

        constexpr bool PULL = sizeof(T) == 4 && MODE == 0;
        Marcher<T, TIN, PULL, TL> m;
        m.init(d, in, b, j0, k0, TJa, TKa, pull);

        const long long rb = (long long)b * d.NI, tb = (long long)b * d.NJ;
        if (LEAN) {
    #pragma unroll
            for (int r = 0; r < 9; r++) {
                const double *src = r == 0 ? t.xi : r == 1 ? t.rho0 : r == 2 ? t.gp : r == 3 ? t.gq : r == 4 ? t.rinv
                                  : r == 5 ? t.r2inv : r == 6 ? t.xir2 : r == 7 ? t.heat : t.g_over_rho0;
                for (int I = threadIdx.x; I < d.NI; I += TL * TL) crow[r][I] = src[rb + I];
            }
            // 1/rho0 for the reciprocal form of the FP32 path's right-hand side
            for (int I = threadIdx.x; I < d.NI; I += TL * TL) crow[9][I] = 1.0 / t.rho0[rb + I];
        }
        const BlockMeta mt = t.meta[b];
        // the right-hand side is computed in FP64 whatever the storage type T:
        // coefficients from the FP64 tables, stage values widened on use
        const double cr = mt.cr, ct = mt.ct, cp = mt.cp;
        // theta rows are read only by active lanes (inactive lanes of a partial
        // tile have J beyond the padded range)
        double st[5];
    #pragma unroll
        for (int q = 0; q < 5; q++) st[q] = m.act ? t.sint[tb + m.J - 2 + q] : 1.0;
        const double sinv = m.act ? t.sinv[tb + m.J] : 1.0, cot = m.act ? t.cot[tb + m.J] : 0.0;
        const double dt = MODE == 0 ? *dtp : 0.0;
        // the FP32 path's reciprocal form (cell_rhs): the stage increment
        // (dt*R)/(5-k) as a product with 1/(5-k)
        constexpr bool RECIP = LEAN;
        const double rdiv = 1.0 / div;

        T win[8][5];
        T nxt[8], hn[8];
    #pragma unroll
        for (int v = 0; v < 8; v++) {
    #pragma unroll
            for (int q = 0; q < 5; q++) win[v][q] = m.col(v, q);
            hn[v] = m.halo(v, d.ng);
        }
        const long long ubase = (long long)b * 8 * d.P + m.coff;
        // G ([block][3][interior cells]) at this lane's column; plane i at + i nj nk
        const long long gnn = (long long)d.ni * d.nj * d.nk;
        const double *gcol = GLM ? glm + (b * 3 * gnn + ((long long)(j0 + ty) * d.nk + (k0 + tx))) : nullptr;

        for (int i = 0; i < d.ni; i++) {
            const int I = i + d.ng;
            const int buf = i & 1;
            // U of this plane, issued first: the plane's arithmetic covers its latency
            TB ub[8];
            if (MODE == 0 && LEAN) {
    #pragma unroll
                for (int v = 0; v < 8; v++) ub[v] = U[ubase + (long long)v * d.P + (long long)I * m.SI];
            }
            // stage this plane: own column centre plus halo strips
    #pragma unroll
            for (int v = 0; v < 8; v++) {
                if (m.act) lds[buf][v][ty + 2][tx + 2] = win[v][2];
                if (m.hv) lds[buf][v][m.hly][m.hlx] = hn[v];
            }
            // prefetch the next window plane and the next halo (on the last plane
            // the clamped loads are discarded)
            const int In = min(I + 3, d.NI - 1), Ih = min(I + 1, d.NI - 1);
            const bool more = i + 1 < d.ni;
    #pragma unroll
            for (int v = 0; v < 8; v++) {
                nxt[v] = (LEAN || (m.act && more)) ? m.col(v, In) : T(0);
                hn[v] = (LEAN || (m.hv && more)) ? m.halo(v, Ih) : T(0);
            }
            __syncthreads();
            if (m.act) {
                // coefficient row r at plane I: LDS (FP32 path) or the table itself
                auto coef = [&](int r, int II) -> double {
                    if (LEAN) return crow[r][II];
                    const double *src = r == 0 ? t.xi : r == 1 ? t.rho0 : r == 2 ? t.gp : r == 3 ? t.gq : r == 4 ? t.rinv
                                      : r == 5 ? t.r2inv : r == 6 ? t.xir2 : r == 7 ? t.heat : t.g_over_rho0;
                    return r == 9 ? 1.0 / t.rho0[rb + II] : src[rb + II];
                };
                double xr[5], rho0r[5], gpr[5], gqr[5];
    #pragma unroll
                for (int q = 0; q < 5; q++) {
                    xr[q] = coef(0, I - 2 + q);
                    rho0r[q] = coef(1, I - 2 + q);
                    gpr[q] = coef(2, I - 2 + q);
                    gqr[q] = coef(3, I - 2 + q);
                }
                auto L = [&](int v, int dj, int dk) -> double { return (double)lds[buf][v][ty + 2 + dj][tx + 2 + dk]; };
                double R[8];
                cell_rhs<double, T, RECIP>(win, L, xr, rho0r, gpr, gqr, st, coef(4, I), coef(5, I), coef(6, I),
                                           coef(7, I), coef(8, I), sinv, cot, cr, ct, cp, RECIP ? coef(9, I) : 1.0, R);
                if (MODE == 0) {
                    if constexpr (GLM && sizeof(T) == 8) {
                        // ModDivB_GLM (3_6/ModDivB.f90:27-28): R(B) = R(B) - grad(psi)
                        const double *gi = gcol + i * (d.nj * d.nk);
                        R[4] = R[4] - gi[0];
                        R[5] = R[5] - gi[gnn];
                        R[6] = R[6] - gi[2 * gnn];
                    }
                    const long long base = (long long)b * 8 * d.P + (long long)I * m.SI + m.coff;
    #pragma unroll
                    for (int v = 0; v < 8; v++) {
                        const long long q = base + (long long)v * d.P;
                        const double inc = RECIP ? (dt * R[v]) * rdiv : (dt * R[v]) / div;
                        const double x = (double)(LEAN ? ub[v] : U[q]) + inc;
                        out[q] = (TOUT)x;
                        if (SHADOW) shadow[q] = (float)x;
                    }
                } else {
                    const long long nn = (long long)d.ni * d.nj * d.nk;
                    const long long base = (long long)b * 8 * nn + ((long long)i * d.nj + (j0 + ty)) * d.nk + (k0 + tx);
    #pragma unroll
                    for (int v = 0; v < 8; v++) ((double *)out)[base + (long long)v * nn] = (double)R[v];
                }
            }
    #pragma unroll
            for (int v = 0; v < 8; v++) {
                win[v][0] = win[v][1]; win[v][1] = win[v][2]; win[v][2] = win[v][3]; win[v][3] = win[v][4];
                win[v][4] = nxt[v];
            }
        }
        if (!PERSIST) break;
        // the next tile's LDS writes wait for this tile's last reads
        if (tile + (int)gridDim.x < ntot) __syncthreads();
    }
}

template <typename T, typename TIN, typename TOUT, typename TB, int MODE, bool SHADOW, int TL, bool PERSIST,
          bool GLM = false>
__global__ void __launch_bounds__(TL * TL) k_stage(Dims d, DevTables t, const TIN *__restrict__ in,
                                               const TB *__restrict__ U, TOUT *__restrict__ out,
                                               float *__restrict__ shadow, const double *__restrict__ dtp,
                                               double div, const int32_t *__restrict__ pull,
                                               const double *__restrict__ glm) {
    stage_body<T, TIN, TOUT, TB, MODE, SHADOW, TL, PERSIST, GLM>(d, t, in, U, out, shadow, dtp, div, pull, glm);
}

// The FP64 step's stage (16-edge tiles, one workgroup per tile) held to two
// wavefronts per SIMD: left to itself the compiler gives it 256 VGPRs plus 2
// AGPRs, one wavefront per SIMD; held to two it spills two values (12 bytes
// per lane) and runs in 3.06 instead of 3.94 ms on the 50 M-cell shell.  The
// other builds either fit two wavefronts already (the FP32 stages) or would
// spill tens of registers.
template <typename T, typename TIN, typename TOUT, typename TB, int MODE, bool SHADOW, int TL, bool PERSIST,
          bool GLM = false>
__global__ void __launch_bounds__(TL * TL) __attribute__((amdgpu_waves_per_eu(2)))
k_stage_w2(Dims d, DevTables t, const TIN *__restrict__ in, const TB *__restrict__ U, TOUT *__restrict__ out,
           float *__restrict__ shadow, const double *__restrict__ dtp, double div,
           const int32_t *__restrict__ pull, const double *__restrict__ glm) {
    stage_body<T, TIN, TOUT, TB, MODE, SHADOW, TL, PERSIST, GLM>(d, t, in, U, out, shadow, dtp, div, pull, glm);
}

// ---------------------------------------------------------------- diffusion

// Radial mirror parities of ModBoundary_Dynamo_primitives (3_6/ModBoundary.f90:
// 16-52) for rho1, vr, vt, vp, br, bt, bp, s1: bottom (inner) and top (outer).
__constant__ double kParBottom[8] = {1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 1.0};
__constant__ double kParTop[8] = {1.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0};

// ModDiffusion_Aritificial_1 (3_6/ModDiffusion.f90:9-150), h = 2, all 8 vars.
// The characteristic speed c of each face is read where the reference reads it:
// the section c(1:ni+1,:,:) (:63) spans the padded transverse range while flux
// spans 1:nj,1:nk, so the speed comes from transverse index (index - ng); same
// for :96 and :130.
// MODE 0: out = U + (0.5*dt)*R (3_6/ModAdvance.f90:104-107); `shadow` (if not
// null) receives the same value in FP32, and the workgroup's min(h/v_wave) of the
// new state (the next step's ModTimeStep_Dynamo input) goes to partial[tile].
// MODE 1: write R.  `in` is the stencil input (the state or its FP32 shadow).
// FP32 path (T = float, MODE 0): U and out are the same buffer, updated in
// place, and Uold is U^n.  The radial ghosts of the input are the mirror of
// U^n that the first stage's boundary call left (the reference applies
// boundaries only to each stage's input), while the interior is the state
// after the fourth stage; at the domain's radial boundary faces the limiter
// therefore divides by the one-step change of the first cell, which FP32
// cannot resolve.  There the face flux is formed in FP64 from U^n (mirrored
// with the parities of ModBoundary_Dynamo_primitives, 3_6/ModBoundary.f90:
// 16-52) and the FP64 state, as the FP64 path forms it.  Elsewhere the FP32
// path's fluxes are computed in FP32: on the 50 M-cell shell FP64 fluxes
// changed the 100-step FP32-vs-FP64 difference only from 5.2e-7 to 5.5e-7
// (largest variable) and cost 0.8 ms per step.

template <typename T, typename TIN, int MODE, int TL, bool PERSIST>
__global__ void __launch_bounds__(TL * TL) k_diffusion(Dims d, DevTables t, const TIN *__restrict__ in,
                                                   const double *U, double *out, float *__restrict__ shadow,
                                                   double *__restrict__ partial, const double *__restrict__ dtp,
                                                   const int32_t *__restrict__ pull, const double *__restrict__ Uold) {
    // one r-plane with its theta/phi halo; the second barrier of each plane
    // orders its reuse, so it is not double-buffered
    constexpr int PL = tile_pl<TL>();
    __shared__ T lds[8][PL][PL];
    // theta and phi face fluxes of the plane, each computed once: fj[v][y][x]
    // is the flux through the lower theta face of tile cell (y, x) (row TJa:
    // the upper face of the last row), fk[v][y][x] through its lower phi face
    __shared__ T fj[8][TL + 1][TL];
    __shared__ T fk[8][TL][TL + 1];
    __shared__ double red[4];
    // block read in place of this block's ghosts in direction code c (the
    // pull table's row, -1 where the ghost itself is read); set before the
    // first barrier of the march
    __shared__ int pdir[27];
    // LEAN (FP32 path, MODE 0): the block's radial rows c_s, xi, sqrt(rho0) are
    // staged in LDS once per tile, and every global load a plane needs besides
    // the next plane's window and halo (U^n of the plane, the shifted speed
    // cells) is issued at the top of the plane, before that prefetch, so no
    // wait in the plane's arithmetic includes the prefetch
    constexpr bool LEAN = sizeof(T) == 4 && MODE == 0;
    __shared__ double drow[3][kMaxNI];
    // LEAN: the theta and phi minmod slopes of the plane's cells, each formed
    // once and read by both faces of the cell: sj[v][y+1][x] for tile rows
    // y = -1 .. TJa, sk[v][y][x+1] for tile columns x = -1 .. TKa
    __shared__ T sj[LEAN ? 8 : 1][TL + 2][TL];
    __shared__ T sk[LEAN ? 8 : 1][TL][TL + 2];
    const int ntk = (d.nk + TL - 1) / TL, tpb = ((d.nj + TL - 1) / TL) * ntk;
    const int ntot = d.nblocks * tpb;
    // PERSIST: a grid smaller than the tile count (a persistent launch) takes
    // the next tile at a stride of the grid size; otherwise one pass, tile =
    // the workgroup
    for (int tile = PERSIST ? (int)blockIdx.x : xcd_tile(blockIdx.x, ntot); tile < ntot; tile += gridDim.x) {
        const int b = tile / tpb;
        if (sizeof(T) == 4 && MODE == 0 && threadIdx.x < 27) pdir[threadIdx.x] = pull ? pull[b * 27 + threadIdx.x] : -1;
        const int trem = tile - b * tpb;
        const int tj = trem / ntk, tk = trem - tj * ntk;
        const int j0 = tj * TL, k0 = tk * TL;
        const int TJa = min(TL, d.nj - j0), TKa = min(TL, d.nk - k0);
        const int tx = threadIdx.x & (TL - 1), ty = threadIdx.x / TL;

        constexpr bool PULL = sizeof(T) == 4 && MODE == 0;
        Marcher<T, TIN, PULL, TL> m;
        m.init(d, in, b, j0, k0, TJa, TKa, pull);
        const long long rb = (long long)b * d.NI, tb = (long long)b * d.NJ;
        if (LEAN) {
            for (int I = threadIdx.x; I < d.NI; I += TL * TL) {
                drow[0][I] = t.cs[rb + I];
                drow[1][I] = t.xi[rb + I];
                drow[2][I] = t.sqrt_rho0[rb + I];
            }
        }
        // radial rows at plane I: LDS (LEAN) or the tables
        auto cs_at = [&](int I) -> double { return LEAN ? drow[0][I] : t.cs[rb + I]; };
        auto xi_at = [&](int I) -> double { return LEAN ? drow[1][I] : t.xi[rb + I]; };
        auto sr_at = [&](int I) -> double { return LEAN ? drow[2][I] : t.sqrt_rho0[rb + I]; };
        const BlockMeta mt = t.meta[b];
        const T dxi = (T)mt.dxi, dxj = (T)mt.dxj, dxk = (T)mt.dxk;
        const T rdxi = LEAN ? (T)__builtin_amdgcn_rcpf((float)dxi) : T(0);
        double hmin = INFINITY;
        const double sint64 = m.act ? t.sint[tb + m.J] : 1.0;
        const T sintc = m.act ? (T)t.sint[tb + m.J] : T(1);
        const double dt = MODE == 0 ? *dtp : 0.0;

        T win[8][5];
        T nxt[8], hn[8];
        // the upper r-face flux of plane i is the lower r-face flux of plane i+1:
        // same four values, same speed sum (speed at the upper cell first), so it
        // is carried instead of recomputed.  The same holds for the theta and phi
        // faces between neighbouring lanes, which go through fj / fk.
        T fr[8];
        T sr[8];   // LEAN: radial slope of the cell above this plane's (carried)
        bool have_fr = false;
        // FP32 path: U before the diffusion on the previous plane (the top
        // boundary face reads it after that plane has been updated in place)
        constexpr bool BFACE = MODE == 0 && sizeof(T) == 4;
        double uprev[8];
        const bool bot = BFACE && (mt.flags & 1), topb = BFACE && (mt.flags & 2);
    #pragma unroll
        for (int v = 0; v < 8; v++) {
    #pragma unroll
            for (int q = 0; q < 5; q++) win[v][q] = (LEAN || m.act) ? m.col(v, q) : T(0);
            hn[v] = (LEAN || m.hv) ? m.halo(v, d.ng) : T(0);
            fr[v] = T(0);
        }
        // a ghost in a direction listed in the pull table is the same-level
        // neighbour's interior cell, translated by n
        auto at = [&](int v, int II, int JJ, int KK) -> T {
            int B = b;
            if (PULL) {
                const int di = II < d.ng ? -1 : (II >= d.ng + d.ni ? 1 : 0);
                const int dj = JJ < d.ng ? -1 : (JJ >= d.ng + d.nj ? 1 : 0);
                const int dk = KK < d.ng ? -1 : (KK >= d.ng + d.nk ? 1 : 0);
                const int D = pdir[(di + 1) * 9 + (dj + 1) * 3 + (dk + 1)];
                if ((di | dj | dk) && D >= 0) {
                    B = D;
                    II -= di * d.ni; JJ -= dj * d.nj; KK -= dk * d.nk;
                }
            }
            return (T)in[((long long)B * 8 + v) * d.P + cell_off(d, II, JJ, KK)];
        };
        // faces on the tile's lower edges: lanes 0 .. TL-1 take the theta face
        // below row 0 in column x = lane, lanes TL .. 2TL-1 the phi face left of
        // column 0 in row y = lane - TL.  Such a face's four-cell stencil lies in
        // the LDS plane at (bstart, step bstep); its speed cells are (bJ0, bK0) on
        // the upper side and (bJ1, bK1) on the lower side.
        const int lane = threadIdx.x;
        const bool bact = lane < TL ? lane < TKa : (lane < 2 * TL && lane - TL < TJa);
        const bool bj = lane < TL;
        const int bx = bj ? lane : 0, by = bj ? 0 : lane - TL;
        const int bstart = bj ? bx + 2 : (by + 2) * PL, bstep = bj ? PL : 1;
        const int bvar = bj ? 2 : 3;
        const int bJ0 = bj ? j0 + d.ng : j0 + by + d.ng - 2, bK0 = bj ? k0 + bx + d.ng - 2 : k0 + d.ng;
        const int bJ1 = bj ? bJ0 - 1 : bJ0, bK1 = bj ? bK0 : bK0 - 1;

        for (int i = 0; i < d.ni; i++) {
            const int I = i + d.ng;
            // LEAN: U^n of the plane and |v_d| at the shifted speed cells, issued
            // before the prefetch (inactive lanes read at the tile's first column)
            double ucur[8];
            T a20 = T(0), a2p = T(0), a30 = T(0), a3p = T(0), a1m = T(0), a10 = T(0), a1p = T(0), as0 = T(0),
              as1 = T(0);
            if (LEAN) {
                const int J = m.act ? m.J : j0 + d.ng, K = m.act ? m.K : k0 + d.ng;
                const long long base = (long long)b * 8 * d.P + (long long)I * m.SI + m.coff;
    #pragma unroll
                for (int v = 0; v < 8; v++) ucur[v] = U[base + (long long)v * d.P];
                a20 = fabs(at(2, I - 2, J, K - 2));
                a2p = fabs(at(2, I - 2, J + 1, K - 2));
                a30 = fabs(at(3, I - 2, J - 2, K));
                a3p = fabs(at(3, I - 2, J - 2, K + 1));
                a1m = fabs(at(1, I - 1, J - 2, K - 2));
                a10 = fabs(at(1, I, J - 2, K - 2));
                a1p = fabs(at(1, I + 1, J - 2, K - 2));
                if (bact) {
                    as0 = fabs(at(bvar, I - 2, bJ0, bK0));
                    as1 = fabs(at(bvar, I - 2, bJ1, bK1));
                }
            }
    #pragma unroll
            for (int v = 0; v < 8; v++) {
                if (m.act) lds[v][ty + 2][tx + 2] = win[v][2];
                if (m.hv) lds[v][m.hly][m.hlx] = hn[v];
            }
            const bool more = i + 1 < d.ni;
            const int In = min(I + 3, d.NI - 1), Ih = min(I + 1, d.NI - 1);
    #pragma unroll
            for (int v = 0; v < 8; v++) {
                nxt[v] = (LEAN || (m.act && more)) ? m.col(v, In) : T(0);
                hn[v] = (LEAN || (m.hv && more)) ? m.halo(v, Ih) : T(0);
            }
            __syncthreads();
            const T csI2 = (T)cs_at(I - 2);
            auto Lv = [&](int v, int dj, int dk) -> T { return lds[v][ty + 2 + dj][tx + 2 + dk]; };
            if (m.act) {
                const int J = m.J, K = m.K;
                // speeds |v_d| + c_s at the cells the reference reads (see above)
                const T c20 = (LEAN ? a20 : fabs(at(2, I - 2, J, K - 2))) + csI2;
                const T c2p = (LEAN ? a2p : fabs(at(2, I - 2, J + 1, K - 2))) + csI2;
                const T c30 = (LEAN ? a30 : fabs(at(3, I - 2, J - 2, K))) + csI2;
                const T c3p = (LEAN ? a3p : fabs(at(3, I - 2, J - 2, K + 1))) + csI2;
                if (LEAN) {
                    // this cell's slopes; its faces are formed after the barrier
    #pragma unroll
                    for (int v = 0; v < 8; v++) {
                        sj[v][ty + 1][tx] = minmod(Lv(v, -1, 0), Lv(v, 0, 0), Lv(v, 1, 0));
                        sk[v][ty][tx + 1] = minmod(Lv(v, 0, -1), Lv(v, 0, 0), Lv(v, 0, 1));
                    }
                } else {
    #pragma unroll
                    for (int v = 0; v < 8; v++) {
                        // upper theta face (the lower one of the cell at row + 1)
                        fj[v][ty + 1][tx] =
                            (face_flux(Lv(v, -1, 0), Lv(v, 0, 0), Lv(v, 1, 0), Lv(v, 2, 0)) * (c2p + c20)) * T(0.5);
                        // upper phi face
                        fk[v][ty][tx + 1] =
                            (face_flux(Lv(v, 0, -1), Lv(v, 0, 0), Lv(v, 0, 1), Lv(v, 0, 2)) * (c3p + c30)) * T(0.5);
                    }
                }
            }
            if (LEAN) {
                // slopes of the halo cells next to the tile: rows -1 and TJa (theta),
                // columns -1 and TKa (phi), lanes 0 .. 4*TL-1
                const int q = lane / TL, x = lane - q * TL;
                if (q < 4 && x < (q < 2 ? TKa : TJa)) {
                    // LDS plane index of the slope cell and the step to its neighbours
                    const int y = q == 0 ? -1 : (q == 1 ? TJa : x), xx = q < 2 ? x : (q == 2 ? -1 : TKa);
                    const int st = q < 2 ? PL : 1;
    #pragma unroll
                    for (int v = 0; v < 8; v++) {
                        const T *pc = &lds[v][y + 2][xx + 2];
                        const T sl = minmod(pc[-st], pc[0], pc[st]);
                        if (q < 2) sj[v][y + 1][x] = sl;
                        else sk[v][x][xx + 1] = sl;
                    }
                }
                __syncthreads();
                if (m.act) {
                    const T c20 = a20 + csI2, c2p = a2p + csI2, c30 = a30 + csI2, c3p = a3p + csI2;
    #pragma unroll
                    for (int v = 0; v < 8; v++) {
                        fj[v][ty + 1][tx] =
                            (face_flux_s<T, true>(Lv(v, 0, 0), Lv(v, 1, 0), sj[v][ty + 1][tx], sj[v][ty + 2][tx]) *
                             (c2p + c20)) * T(0.5);
                        fk[v][ty][tx + 1] =
                            (face_flux_s<T, true>(Lv(v, 0, 0), Lv(v, 0, 1), sk[v][ty][tx + 1], sk[v][ty][tx + 2]) *
                             (c3p + c30)) * T(0.5);
                    }
                }
            }
            if (bact) {
                // the lower face of the edge cell
                const T s0 = (LEAN ? as0 : fabs(at(bvar, I - 2, bJ0, bK0))) + csI2;
                const T s1 = (LEAN ? as1 : fabs(at(bvar, I - 2, bJ1, bK1))) + csI2;
    #pragma unroll
                for (int v = 0; v < 8; v++) {
                    const T *pv = &lds[v][0][0] + bstart;
                    T f;
                    if (LEAN)
                        f = face_flux_s<T, true>(pv[bstep], pv[2 * bstep], bj ? sj[v][0][bx] : sk[v][by][0],
                                                 bj ? sj[v][1][bx] : sk[v][by][1]);
                    else
                        f = face_flux(pv[0], pv[bstep], pv[2 * bstep], pv[3 * bstep]);
                    f = (f * (s0 + s1)) * T(0.5);
                    if (bj) fj[v][0][bx] = f;
                    else fk[v][by][0] = f;
                }
            }
            __syncthreads();
            if (m.act) {
                const int J = m.J, K = m.K;
                const T csI = (T)cs_at(I), csIm = (T)cs_at(I - 1), csIp = (T)cs_at(I + 1);
                const T c1m = (LEAN ? a1m : fabs(at(1, I - 1, J - 2, K - 2))) + csIm;
                const T c10 = (LEAN ? a10 : fabs(at(1, I, J - 2, K - 2))) + csI;
                const T c1p = (LEAN ? a1p : fabs(at(1, I + 1, J - 2, K - 2))) + csIp;
                const T xic = (T)xi_at(I);
                const T L2 = dxj * xic;
                const T L3 = (dxk * xic) * sintc;
                // LEAN: the face sums scaled by hardware reciprocals of the lengths
                const T rL2 = LEAN ? (T)__builtin_amdgcn_rcpf((float)L2) : T(0);
                const T rL3 = LEAN ? (T)__builtin_amdgcn_rcpf((float)L3) : T(0);
                const long long base = (long long)b * 8 * d.P + (long long)I * m.SI + m.coff;
                const long long nn = (long long)d.ni * d.nj * d.nk;
                const long long rbase = (long long)b * 8 * nn + ((long long)i * d.nj + (j0 + ty)) * d.nk + (k0 + tx);
                double vv = 0.0, bb2 = 0.0;   // |v|^2 and |B|^2 of the new state, summed in component order
    #pragma unroll
                for (int v = 0; v < 8; v++) {
                    T r = T(0);
                    T f0, f1;
                    if (LEAN) {
                        // radial slopes of this cell and the one above; the latter is
                        // carried to the next plane
                        const T sc = have_fr ? sr[v] : minmod(win[v][1], win[v][2], win[v][3]);
                        const T su = minmod(win[v][2], win[v][3], win[v][4]);
                        f0 = have_fr ? fr[v]
                                     : (face_flux_s<T, true>(win[v][1], win[v][2],
                                                             minmod(win[v][0], win[v][1], win[v][2]), sc) *
                                        (c10 + c1m)) * T(0.5);
                        f1 = (face_flux_s<T, true>(win[v][2], win[v][3], sc, su) * (c1p + c10)) * T(0.5);
                        sr[v] = su;
                    } else {
                        f0 = have_fr ? fr[v]
                                     : (face_flux(win[v][0], win[v][1], win[v][2], win[v][3]) * (c10 + c1m)) * T(0.5);
                        f1 = (face_flux(win[v][1], win[v][2], win[v][3], win[v][4]) * (c1p + c10)) * T(0.5);
                    }
                    if (BFACE && bot && i == 0) {
                        // lower face of the first cell: ghosts 0 and -1 mirror cells 1 and 2 of U^n
                        const long long q = base + (long long)v * d.P;
                        const double p = kParBottom[v];
                        const double g0 = p * Uold[q], gm1 = p * Uold[q + m.SI];
                        f0 = (T)((face_flux<double>(gm1, g0, ucur[v], U[q + m.SI]) * (double)(c10 + c1m)) * 0.5);
                    }
                    if (BFACE && topb && i == d.ni - 1) {
                        // upper face of the last cell: ghosts ni+1 and ni+2 mirror cells ni and ni-1 of U^n
                        const long long q = base + (long long)v * d.P;
                        const double p = kParTop[v];
                        const double g1 = p * Uold[q], g2 = p * Uold[q - m.SI];
                        f1 = (T)((face_flux<double>(uprev[v], ucur[v], g1, g2) * (double)(c1p + c10)) * 0.5);
                    }
                    fr[v] = f1;
                    if (LEAN) {
                        r = r + (f0 - f1) * rdxi;
                        r = r + (fj[v][ty][tx] - fj[v][ty + 1][tx]) * rL2;
                        r = r + (fk[v][ty][tx] - fk[v][ty][tx + 1]) * rL3;
                    } else {
                        r = r + (f0 - f1) / dxi;
                        r = r + (fj[v][ty][tx] - fj[v][ty + 1][tx]) / L2;
                        r = r + (fk[v][ty][tx] - fk[v][ty][tx + 1]) / L3;
                    }
                    if (MODE == 0) {
                        const long long q = base + (long long)v * d.P;
                        const double x = (LEAN ? ucur[v] : U[q]) + (0.5 * dt) * (double)r;
                        out[q] = x;
                        if (shadow) shadow[q] = (float)x;
                        if (v == 1) vv = x * x;
                        if (v == 2 || v == 3) vv = vv + x * x;
                        if (v == 4) bb2 = x * x;
                        if (v == 5 || v == 6) bb2 = bb2 + x * x;
                    } else {
                        out[rbase + (long long)v * nn] = (double)r;
                    }
                }
                if (MODE == 0) {
                    // ModWaveSpeed_Dynamo (3_6/ModWaveSpeed.f90:12-24) on the new state and
                    // h_LLL (3_6/ModBlock.f90:420-422), as k_dt_tile computes them
                    double w = cs_at(I) + sqrt(vv);
                    w = w + sqrt(bb2) / sr_at(I);
                    const double xi = xi_at(I);
                    double h = mt.dxi, h2 = xi * mt.dxj, h3 = (xi * sint64) * mt.dxk;
                    h = h < h2 ? h : h2;
                    h = h < h3 ? h : h3;
                    const double x = h / w;
                    hmin = x < hmin ? x : hmin;
                }
            }
            have_fr = true;
    #pragma unroll
            for (int v = 0; v < 8; v++) {
                win[v][0] = win[v][1]; win[v][1] = win[v][2]; win[v][2] = win[v][3]; win[v][3] = win[v][4];
                win[v][4] = nxt[v];
                if (BFACE) uprev[v] = ucur[v];
            }
        }
        if (MODE == 0 && partial) {
            // min is exact in any order: wavefront shuffle, then across the wavefronts
            for (int off = 32; off > 0; off >>= 1) {
                double o = __shfl_down(hmin, off, 64);
                hmin = o < hmin ? o : hmin;
            }
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = hmin;
            __syncthreads();
            if (threadIdx.x == 0) {
                double r = red[0];
                for (int q = 1; q < TL * TL / 64; q++) r = red[q] < r ? red[q] : r;
                partial[tile] = r;
            }
        }
        if (!PERSIST) break;
        // the next tile's LDS writes wait for this tile's last reads
        if (tile + (int)gridDim.x < ntot) __syncthreads();
    }
}

// ---------------------------------------------------------------- exchange

// ModCommunication_SendRecvGC_new (3_6/ModCommunication.f90:678-721) as one
// table-driven gather: trilinear in the fixed term order of ModMath.f90:283-291
// with FP64 weights and FP64 accumulation, then the Yin-Yang rotation of v and B
// (ModYinYang_VecConv_1D, ModYinYang.f90:59-60).  Each ghost has one donor, and
// donors are read only at interior cells (clamped bases touch a donor ghost only
// with an exactly-zero weight), so entries are independent.
//
// SKIP: a corner whose weight factor is exactly zero contributes f*0 = +-0 to
// the sum in the reference; its load is skipped and the term taken as +0, which
// leaves every nonzero result bit-identical.  On the production layouts this
// removes 60% of the FP64 corner loads and shortens the FP64 fill by 15%.  For
// the FP32 buffers (whose near-copies go to the copy table) the unpredicated
// gather measures twice as fast.  A zero-weight corner may then lie on a donor
// ghost cell that the ghost pull leaves unwritten; the FP32 buffers are zeroed
// when allocated and written with finite values only, so that term is f*0 = +-0
// as well.
// PSI: also psi of GLM divergence cleaning ([block][P], FP64), a scalar
template <typename TS, bool SKIP, bool PSI = false>
__global__ void __launch_bounds__(256) k_exchange(Dims d, DevGhosts g, TS *__restrict__ st,
                                                  double *__restrict__ psi = nullptr) {
    const int e = g.first + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= g.n_total) return;
    const long long P = d.P;
    // the base cell is interior along phi, so its corners are in the main part
    const int SI = (int)d.SI, SJ = d.SJ;
    const int dst = g.dst[e], src = g.src[e];
    const double wi = g.w[3 * (long long)e], wj = g.w[3 * (long long)e + 1], wk = g.w[3 * (long long)e + 2];
    const bool i0 = !SKIP || wi != 1.0, i1 = !SKIP || wi != 0.0, j0 = !SKIP || wj != 1.0, j1 = !SKIP || wj != 0.0,
               k0 = !SKIP || wk != 1.0, k1 = !SKIP || wk != 0.0;
    double val[8];
#pragma unroll
    for (int v = 0; v < 8; v++) {
        const TS *F = st + src + (long long)v * P;
        double f000, f100, f010, f001, f110, f011, f101, f111;
        if constexpr (sizeof(TS) == 8 && SKIP) {
            // the two phi corners of each (i, j) pair are adjacent: one 16-byte
            // load per pair (the phi neighbour of a clamped base at the last
            // interior cell is read in bounds and dropped with its zero weight)
            auto ld2 = [](const TS *p) -> double2 {
                double2 r;
                __builtin_memcpy(&r, p, sizeof(r));
                return r;
            };
            const double2 z = make_double2(0.0, 0.0);
            const double2 a00 = (i0 && j0) ? ld2(F) : z, a10 = (i1 && j0) ? ld2(F + SI) : z;
            const double2 a01 = (i0 && j1) ? ld2(F + SJ) : z, a11 = (i1 && j1) ? ld2(F + SI + SJ) : z;
            f000 = k0 ? a00.x : 0.0; f001 = k1 ? a00.y : 0.0;
            f100 = k0 ? a10.x : 0.0; f101 = k1 ? a10.y : 0.0;
            f010 = k0 ? a01.x : 0.0; f011 = k1 ? a01.y : 0.0;
            f110 = k0 ? a11.x : 0.0; f111 = k1 ? a11.y : 0.0;
        } else {
            f000 = (i0 && j0 && k0) ? (double)F[0] : 0.0;
            f100 = (i1 && j0 && k0) ? (double)F[SI] : 0.0;
            f010 = (i0 && j1 && k0) ? (double)F[SJ] : 0.0;
            f001 = (i0 && j0 && k1) ? (double)F[1] : 0.0;
            f110 = (i1 && j1 && k0) ? (double)F[SI + SJ] : 0.0;
            f011 = (i0 && j1 && k1) ? (double)F[SJ + 1] : 0.0;
            f101 = (i1 && j0 && k1) ? (double)F[SI + 1] : 0.0;
            f111 = (i1 && j1 && k1) ? (double)F[SI + SJ + 1] : 0.0;
        }
        double t = f000 * (1.0 - wi) * (1.0 - wj) * (1.0 - wk);
        t = t + f100 * (wi) * (1.0 - wj) * (1.0 - wk);
        t = t + f010 * (1.0 - wi) * (wj) * (1.0 - wk);
        t = t + f001 * (1.0 - wi) * (1.0 - wj) * (wk);
        t = t + f110 * (wi) * (wj) * (1.0 - wk);
        t = t + f011 * (1.0 - wi) * (wj) * (wk);
        t = t + f101 * (wi) * (1.0 - wj) * (wk);
        t = t + f111 * (wi) * (wj) * (wk);
        val[v] = t;
    }
    if (e >= g.yy_begin) {
        const long long q = 2 * (long long)e;
        const double a = g.rot[q], bb = g.rot[q + 1];
        double b2 = -(a * val[5]) - bb * val[6];
        double b3 = -(a * val[6]) + bb * val[5];
        val[5] = b2; val[6] = b3;
        double v2 = -(a * val[2]) - bb * val[3];
        double v3 = -(a * val[3]) + bb * val[2];
        val[2] = v2; val[3] = v3;
    }
#pragma unroll
    for (int v = 0; v < 8; v++) st[dst + (long long)v * P] = (TS)val[v];
    if constexpr (PSI) {
        // psi has the state's cell layout: the same paired phi-corner loads
        const double *F = psi + psi_off(d, src);
        auto ld2 = [](const double *p) -> double2 {
            double2 r;
            __builtin_memcpy(&r, p, sizeof(r));
            return r;
        };
        const double2 z = make_double2(0.0, 0.0);
        const double2 a00 = (i0 && j0) ? ld2(F) : z, a10 = (i1 && j0) ? ld2(F + SI) : z;
        const double2 a01 = (i0 && j1) ? ld2(F + SJ) : z, a11 = (i1 && j1) ? ld2(F + SI + SJ) : z;
        const double f000 = k0 ? a00.x : 0.0, f001 = k1 ? a00.y : 0.0;
        const double f100 = k0 ? a10.x : 0.0, f101 = k1 ? a10.y : 0.0;
        const double f010 = k0 ? a01.x : 0.0, f011 = k1 ? a01.y : 0.0;
        const double f110 = k0 ? a11.x : 0.0, f111 = k1 ? a11.y : 0.0;
        double t = f000 * (1.0 - wi) * (1.0 - wj) * (1.0 - wk);
        t = t + f100 * (wi) * (1.0 - wj) * (1.0 - wk);
        t = t + f010 * (1.0 - wi) * (wj) * (1.0 - wk);
        t = t + f001 * (1.0 - wi) * (1.0 - wj) * (wk);
        t = t + f110 * (wi) * (wj) * (1.0 - wk);
        t = t + f011 * (1.0 - wi) * (wj) * (wk);
        t = t + f101 * (wi) * (1.0 - wj) * (wk);
        t = t + f111 * (wi) * (wj) * (wk);
        psi[psi_off(d, dst)] = t;
    }
}

// Ghost entries filled by a plain copy, sorted by destination so consecutive
// lanes copy consecutive phi cells.  For FP64 buffers these are the same-panel
// entries whose three weights are exactly 0: the weighted sum reduces to the
// base cell (an interior donor cell, the base being clamped to [1, n]) plus +0
// from the vanishing terms.  For the FP32 buffers of the deviation path they
// also include entries whose weights lie within 1e-12 of 0 or 1 (same-level
// neighbours whose cell centres coincide up to rounding), copied from the
// nearest donor cell.
template <typename TS, bool PSI = false>
__global__ void __launch_bounds__(256) k_exchange_copy(Dims d, const int2 *__restrict__ copy, int n,
                                                       TS *__restrict__ st, double *__restrict__ psi = nullptr) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const int2 c = copy[e];
    const long long P = d.P;
    TS val[8];
#pragma unroll
    for (int v = 0; v < 8; v++) val[v] = st[c.y + (long long)v * P];
#pragma unroll
    for (int v = 0; v < 8; v++) st[c.x + (long long)v * P] = val[v] + TS(0);
    if constexpr (PSI) psi[psi_off(d, c.x)] = psi[psi_off(d, c.y)] + 0.0;
}

// FP32 buffers: same-panel entries whose theta/phi weights lie within 1e-12 of
// 0 or 1 and whose r weight does not (neighbours across a radial block face,
// whose radial spacings differ): the two r corners of the weighted sum.
template <typename TS>
__global__ void __launch_bounds__(256) k_exchange_lerp(Dims d, DevGhosts g, TS *__restrict__ st) {
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= g.n_lerp) return;
    const int2 c = g.lerp[e];
    const double wi = g.lerp_w[e];
    const long long P = d.P;
    const int SI = (int)d.SI;   // the donor cell is interior along phi (main part)
    double val[8];
#pragma unroll
    for (int v = 0; v < 8; v++) {
        const TS *F = st + c.y + (long long)v * P;
        val[v] = (double)F[0] * (1.0 - wi) + (double)F[SI] * wi;
    }
#pragma unroll
    for (int v = 0; v < 8; v++) st[c.x + (long long)v * P] = (TS)val[v];
}

// ---------------------------------------------------------------- boundary

// ModBoundary_Dynamo_primitives (3_6/ModBoundary.f90:10-55): radial mirror of
// the ng ghost planes of top/bottom blocks over all (j,k), parities per variable.
// The mirror source of the theta/phi ghost columns is a ghost cell of the
// block's own interior plane; with a pull table (see Marcher) those in a listed
// direction are read from the same-level neighbour's interior.
template <typename TS>
__global__ void __launch_bounds__(256) k_boundary(Dims d, DevBoundary bl, TS *__restrict__ st,
                                                  const int32_t *__restrict__ pull) {
    const long long plane = (long long)d.NJ * d.NK;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)d.ng * plane;
    if (tid >= per * bl.n) return;
    const int e = (int)(tid / per);
    const long long rem = tid - (long long)e * per;
    const int g = (int)(rem / plane);
    const long long jk = rem - (long long)g * plane;
    const int b = bl.list[e] >> 1, side = bl.list[e] & 1;
    int dI, sI;
    if (side) { dI = d.ni + d.ng + g; sI = d.ni + d.ng - 1 - g; }
    else { dI = g; sI = 2 * d.ng - 1 - g; }
    // signs: top (+,-,+,+,+,-,-,+), bottom (+,-,+,+,-,+,+,+)
    const unsigned neg = side ? 0x62u : 0x12u;   // bit v set -> negate variable v
    // (J, K) of this thread's cell, enumerated row by row over the padded plane
    const int J = (int)(jk / d.NK), K = (int)(jk - (long long)J * d.NK);
    TS *base = st + (long long)b * 8 * d.P;
    const TS *src = base;
    long long soff = cell_off(d, sI, J, K);
    if (pull) {
        const int dj = J < d.ng ? -1 : (J >= d.ng + d.nj ? 1 : 0);
        const int dk = K < d.ng ? -1 : (K >= d.ng + d.nk ? 1 : 0);
        if (dj | dk) {
            const int D = pull[b * 27 + 9 + (dj + 1) * 3 + (dk + 1)];
            if (D >= 0) {
                src = st + (long long)D * 8 * d.P;
                soff = cell_off(d, sI, J - dj * d.nj, K - dk * d.nk);
            }
        }
    }
    const long long doff = cell_off(d, dI, J, K);
#pragma unroll
    for (int v = 0; v < 8; v++) {
        TS x = src[(long long)v * d.P + soff];
        base[(long long)v * d.P + doff] = ((neg >> v) & 1u) ? -x : x;
    }
}

// ---------------------------------------------------------------- dt

// ModWaveSpeed_Dynamo (3_6/ModWaveSpeed.f90:12-24) and minval(h/v_wave) of
// ModTimeStep_Dynamo (3_6/ModTimeStep.f90:27) over one 16x16 (theta,phi) tile
// column of a block, all r.  Used when the state was set from outside; during
// stepping the diffusion kernel produces the same partials.
template <int TL>
__global__ void __launch_bounds__(TL * TL) k_dt_tile(Dims d, DevTables t, const double *__restrict__ U,
                                                     double *__restrict__ partial) {
    __shared__ double red[4];
    const int ntk = (d.nk + TL - 1) / TL, tpb = ((d.nj + TL - 1) / TL) * ntk;
    const int b = blockIdx.x / tpb;
    const int trem = blockIdx.x - b * tpb;
    const int tj = trem / ntk, tk = trem - tj * ntk;
    const int tx = threadIdx.x & (TL - 1), ty = threadIdx.x / TL;
    const int j = tj * TL + ty, k = tk * TL + tx;
    const long long rb = (long long)b * d.NI, tb = (long long)b * d.NJ;
    const BlockMeta mt = t.meta[b];
    const double *Ub = U + (long long)b * 8 * d.P;
    double mn = INFINITY;
    if (j < d.nj && k < d.nk) {
        const int J = j + d.ng, K = k + d.ng;
        const double st = t.sint[tb + J];
        for (int i = 0; i < d.ni; i++) {
            const int I = i + d.ng;
            const long long q = cell_off(d, I, J, K);
            const double vr = Ub[q + d.P], vt = Ub[q + 2 * d.P], vp = Ub[q + 3 * d.P];
            const double br = Ub[q + 4 * d.P], bt = Ub[q + 5 * d.P], bp = Ub[q + 6 * d.P];
            double w = t.cs[rb + I] + sqrt(vr * vr + vt * vt + vp * vp);
            w = w + sqrt(br * br + bt * bt + bp * bp) / t.sqrt_rho0[rb + I];
            const double xi = t.xi[rb + I];
            double h = mt.dxi, h2 = xi * mt.dxj, h3 = (xi * st) * mt.dxk;
            h = h < h2 ? h : h2;
            h = h < h3 ? h : h3;
            const double x = h / w;
            mn = x < mn ? x : mn;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_down(mn, off, 64);
        mn = o < mn ? o : mn;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int q = 1; q < TL * TL / 64; q++) r = red[q] < r ? red[q] : r;
        partial[blockIdx.x] = r;
    }
}

__global__ void __launch_bounds__(256) k_dt_final(const double *__restrict__ partial, int n, double cfl,
                                                  double *__restrict__ dtp) {
    double mn = INFINITY;
    for (int i = threadIdx.x; i < n; i += blockDim.x) mn = partial[i] < mn ? partial[i] : mn;
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_down(mn, off, 64);
        mn = o < mn ? o : mn;
    }
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = red[w] < r ? red[w] : r;
        // dt = min(1.e20, CFL*min) (3_6/ModTimeStep.f90:254-262)
        double dt = cfl * r;
        const double big = (double)1.0e20f;
        dtp[0] = dt < big ? dt : big;
    }
}

// Radial ghost planes of top/bottom blocks copied from one buffer to another.
template <typename TS>
__global__ void __launch_bounds__(256) k_copy_radial(Dims d, DevBoundary bl, const TS *__restrict__ src,
                                                     TS *__restrict__ dst) {
    const long long plane = (long long)d.NJ * d.NK;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)d.ng * plane;
    if (tid >= per * bl.n) return;
    const int e = (int)(tid / per);
    const long long rem = tid - (long long)e * per;
    const int g = (int)(rem / plane);
    const long long jk = rem - (long long)g * plane;
    const int b = bl.list[e] >> 1, side = bl.list[e] & 1;
    const int I = side ? d.ni + d.ng + g : g;
    const int J = (int)(jk / d.NK), K = (int)(jk - (long long)J * d.NK);
    const long long base = (long long)b * 8 * d.P + cell_off(d, I, J, K);
#pragma unroll
    for (int v = 0; v < 8; v++) dst[base + (long long)v * d.P] = src[base + (long long)v * d.P];
}

__global__ void k_to_f32(const double *__restrict__ s, float *__restrict__ d, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = (float)s[i];
}

__global__ void k_set_scalar(double *__restrict__ p, double v) {
    if (threadIdx.x == 0) p[0] = v;
}

// Interior difference of two states of one layout, one workgroup per (block,
// variable).  PASS 0: w[v] = max|a-b|, w[8+v] = max|b| (as the bit patterns of
// non-negative doubles, which order like the values).  PASS 1: w[16+v] = cells
// with |a-b| > tol*max|b|, w[24+v] = the smallest interior index
// ((b*ni + i)*nj + j)*nk + k holding max|a-b|.
template <int PASS>
__global__ void __launch_bounds__(256) k_state_diff(Dims d, const double *__restrict__ a, const double *__restrict__ b,
                                                    double tol, unsigned long long *__restrict__ w) {
    const int blk = blockIdx.x >> 3, v = blockIdx.x & 7;
    const long long n3 = (long long)d.ni * d.nj * d.nk;
    const long long base = ((long long)blk * 8 + v) * d.P;
    double m0 = 0.0, m1 = 0.0;
    unsigned long long cnt = 0, first = ~0ull;
    const double dmax = PASS == 1 ? __longlong_as_double((long long)w[v]) : 0.0;
    const double lim = PASS == 1 ? tol * __longlong_as_double((long long)w[8 + v]) : 0.0;
    for (long long q = threadIdx.x; q < n3; q += blockDim.x) {
        const int k = (int)(q % d.nk), j = (int)((q / d.nk) % d.nj), i = (int)(q / ((long long)d.nk * d.nj));
        const long long o = base + cell_off(d, i + d.ng, j + d.ng, k + d.ng);
        const double x = fabs(a[o] - b[o]), y = fabs(b[o]);
        if (PASS == 0) {
            m0 = x > m0 ? x : m0;
            m1 = y > m1 ? y : m1;
        } else {
            cnt += x > lim ? 1 : 0;
            if (x == dmax) {
                const unsigned long long id = (unsigned long long)((((long long)blk * d.ni + i) * d.nj + j) * d.nk + k);
                first = id < first ? id : first;
            }
        }
    }
    if (PASS == 0) {
        atomicMax(&w[v], (unsigned long long)__double_as_longlong(m0));
        atomicMax(&w[8 + v], (unsigned long long)__double_as_longlong(m1));
    } else {
        if (cnt) atomicAdd(&w[16 + v], cnt);
        if (first != ~0ull) atomicMin(&w[24 + v], first);
    }
}

// ModCheck_primitive (3_6/ModCheck.f90:12-71) on a padded state of any
// precision: w[0] += NaN count, w[1] = min over NaN cells of the reference's
// scan position ((b*NK + K)*NJ + J)*NI*8 + I*8 + v (blocks in order, then
// phi, theta, r, variable innermost), one workgroup per (block, variable).
template <typename TS>
__global__ void __launch_bounds__(256) k_nan_scan(Dims d, const TS *__restrict__ st, unsigned long long *__restrict__ w) {
    const int blk = blockIdx.x >> 3, v = blockIdx.x & 7;
    const TS *p = st + ((long long)blk * 8 + v) * d.P;
    unsigned long long cnt = 0, first = ~0ull;
    for (long long q = threadIdx.x; q < d.P; q += blockDim.x) {
        const TS x = p[q];
        if (x != x) {
            int I, J, K;
            cell_ijk(d, q, I, J, K);
            // variable innermost, with room for psi (variable 8) after the eight
            const unsigned long long key =
                ((((unsigned long long)blk * d.NK + K) * d.NJ + J) * d.NI + I) * 9ull + (unsigned long long)v;
            cnt++;
            first = key < first ? key : first;
        }
    }
    if (cnt) {
        atomicAdd(&w[0], cnt);
        atomicMin(&w[1], first);
    }
}

// psi of GLM divergence cleaning ([block][P]), as the ninth variable of the
// same scan order
__global__ void __launch_bounds__(256) k_nan_scan_psi(Dims d, const double *__restrict__ psi,
                                                      unsigned long long *__restrict__ w) {
    const int blk = blockIdx.x;
    const double *p = psi + (long long)blk * d.P;
    unsigned long long cnt = 0, first = ~0ull;
    for (long long q = threadIdx.x; q < d.P; q += blockDim.x) {
        const double x = p[q];
        if (x != x) {
            int I, J, K;
            cell_ijk(d, q, I, J, K);
            const unsigned long long key = ((((unsigned long long)blk * d.NK + K) * d.NJ + J) * d.NI + I) * 9ull + 8ull;
            cnt++;
            first = key < first ? key : first;
        }
    }
    if (cnt) {
        atomicAdd(&w[0], cnt);
        atomicMin(&w[1], first);
    }
}

// ---------------------------------------------------------------- launchers

// workgroups of a march over every block with tile edge TL
static inline unsigned tiles(const Dims &d, int TL) {
    return (unsigned)((long long)d.nblocks * ((d.nj + TL - 1) / TL) * ((d.nk + TL - 1) / TL));
}

static void check_tile(int tile) {
    if (tile != 8 && tile != 16) throw std::runtime_error("tile edge must be 8 or 16");
}

static void check_stage_ni(const Dims &d) {
    if (d.NI > kMaxNI) throw std::runtime_error("stage march: more than 128 radial cells per block");
}

// grid of a march: one workgroup per tile (wgs = 0) or wgs workgroups that
// loop over the tiles (a persistent launch), never more than the tiles
static inline dim3 march_grid(const Dims &d, int TL, int wgs) {
    if (wgs < 0) throw std::runtime_error("workgroup count < 0");
    const unsigned n = tiles(d, TL);
    return dim3(wgs == 0 ? n : std::min<unsigned>(n, (unsigned)wgs));
}

template <typename K>
static int resident(K kernel, int threads) {
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess)
        throw std::runtime_error("occupancy query failed");
    return std::max(1, per_cu) * prop.multiProcessorCount;
}

// workgroups of one march resident on the device at once (the persistent grid):
// kernel 0 the stage, 1 the diffusion; precision 0 FP64, 1 FP32.  The FP32
// path launches the stage in two builds (stages 1-3 and the final stage with
// its FP64 output and FP32 copy) on one grid size: the smaller of their counts.
int march_resident_wgs(int kernel, int precision, int tile) {
    check_tile(tile);
    if (kernel == 0 && precision == 0)
        return tile == 16 ? resident(k_stage<double, double, double, double, 0, false, 16, true>, 256)
                          : resident(k_stage<double, double, double, double, 0, false, 8, true>, 64);
    if (kernel == 0)
        return tile == 16 ? std::min(resident(k_stage<float, float, float, float, 0, false, 16, true>, 256),
                                     resident(k_stage<float, float, double, double, 0, true, 16, true>, 256))
                          : std::min(resident(k_stage<float, float, float, float, 0, false, 8, true>, 64),
                                     resident(k_stage<float, float, double, double, 0, true, 8, true>, 64));
    if (precision == 0)
        return tile == 16 ? resident(k_diffusion<double, double, 0, 16, true>, 256)
                          : resident(k_diffusion<double, double, 0, 8, true>, 64);
    return tile == 16 ? resident(k_diffusion<float, float, 0, 16, true>, 256)
                      : resident(k_diffusion<float, float, 0, 8, true>, 64);
}

template <int TL, typename T, typename TIN, typename TOUT, typename TB>
static void stage_tl(const Dims &d, const DevTables &t, const TIN *in, const TB *U, TOUT *out, float *shadow,
                     const double *dtp, double div, hipStream_t s, const int32_t *pull, int wgs, const double *glm) {
    const dim3 g = march_grid(d, TL, wgs), b(TL * TL);
    if constexpr (sizeof(T) == 8) {
        if (glm) {
            if (shadow) throw std::runtime_error("GLM stage: no FP32 copy");
            if (wgs)
                hipLaunchKernelGGL((k_stage<T, TIN, TOUT, TB, 0, false, TL, true, true>), g, b, 0, s, d, t, in, U, out,
                                   (float *)nullptr, dtp, div, pull, glm);
            else if (TL == 16)
                hipLaunchKernelGGL((k_stage_w2<T, TIN, TOUT, TB, 0, false, TL, false, true>), g, b, 0, s, d, t, in, U,
                                   out, (float *)nullptr, dtp, div, pull, glm);
            else
                hipLaunchKernelGGL((k_stage<T, TIN, TOUT, TB, 0, false, TL, false, true>), g, b, 0, s, d, t, in, U, out,
                                   (float *)nullptr, dtp, div, pull, glm);
            return;
        }
        if (!shadow && !wgs && TL == 16) {
            hipLaunchKernelGGL((k_stage_w2<T, TIN, TOUT, TB, 0, false, TL, false>), g, b, 0, s, d, t, in, U, out,
                               (float *)nullptr, dtp, div, pull, glm);
            return;
        }
    }
    if (shadow && wgs)
        hipLaunchKernelGGL((k_stage<T, TIN, TOUT, TB, 0, true, TL, true>), g, b, 0, s, d, t, in, U, out, shadow, dtp, div,
                           pull, glm);
    else if (shadow)
        hipLaunchKernelGGL((k_stage<T, TIN, TOUT, TB, 0, true, TL, false>), g, b, 0, s, d, t, in, U, out, shadow, dtp, div,
                           pull, glm);
    else if (wgs)
        hipLaunchKernelGGL((k_stage<T, TIN, TOUT, TB, 0, false, TL, true>), g, b, 0, s, d, t, in, U, out,
                           (float *)nullptr, dtp, div, pull, glm);
    else
        hipLaunchKernelGGL((k_stage<T, TIN, TOUT, TB, 0, false, TL, false>), g, b, 0, s, d, t, in, U, out,
                           (float *)nullptr, dtp, div, pull, glm);
}

template <typename T, typename TIN, typename TOUT, typename TB>
void launch_stage(const Dims &d, const DevTables &t, const TIN *in, const TB *U, TOUT *out, float *shadow,
                  const double *dtp, double div, hipStream_t s, const int32_t *pull, int tile, int wgs,
                  const double *glm) {
    check_tile(tile);
    check_stage_ni(d);
    if (glm && sizeof(T) != 8) throw std::runtime_error("GLM divergence cleaning runs on the FP64 path only");
    if (tile == 16) stage_tl<16, T>(d, t, in, U, out, shadow, dtp, div, s, pull, wgs, glm);
    else stage_tl<8, T>(d, t, in, U, out, shadow, dtp, div, s, pull, wgs, glm);
}

template <int TL, bool PERSIST, typename T, typename TIN>
static void rhs_tl(const Dims &d, const DevTables &t, const TIN *in, double *R, hipStream_t s, int wgs) {
    hipLaunchKernelGGL((k_stage<T, TIN, double, double, 1, false, TL, PERSIST>), march_grid(d, TL, wgs), dim3(TL * TL), 0,
                       s, d, t, in, (const double *)nullptr, R, (float *)nullptr, (const double *)nullptr, 1.0,
                       (const int32_t *)nullptr, (const double *)nullptr);
}

template <typename T, typename TIN>
void launch_rhs_only(const Dims &d, const DevTables &t, const TIN *in, double *R, hipStream_t s, int tile, int wgs) {
    check_tile(tile);
    check_stage_ni(d);
    if (tile == 16) {
        if (wgs) rhs_tl<16, true, T>(d, t, in, R, s, wgs);
        else rhs_tl<16, false, T>(d, t, in, R, s, wgs);
    } else {
        if (wgs) rhs_tl<8, true, T>(d, t, in, R, s, wgs);
        else rhs_tl<8, false, T>(d, t, in, R, s, wgs);
    }
}

template <int TL, bool PERSIST, int MODE, typename T, typename TIN>
static void diffusion_tl(const Dims &d, const DevTables &t, const TIN *in, const double *U, double *out,
                         float *shadow, double *partial, const double *dtp, hipStream_t s, const int32_t *pull,
                         int wgs, const double *Uold) {
    hipLaunchKernelGGL((k_diffusion<T, TIN, MODE, TL, PERSIST>), march_grid(d, TL, wgs), dim3(TL * TL), 0, s, d, t, in,
                       U, out, shadow, partial, dtp, pull, Uold);
}

template <int MODE, typename T, typename TIN>
static void diffusion_any(const Dims &d, const DevTables &t, const TIN *in, const double *U, double *out,
                          float *shadow, double *partial, const double *dtp, hipStream_t s, const int32_t *pull,
                          int tile, int wgs, const double *Uold) {
    check_tile(tile);
    if (tile == 16) {
        if (wgs) diffusion_tl<16, true, MODE, T>(d, t, in, U, out, shadow, partial, dtp, s, pull, wgs, Uold);
        else diffusion_tl<16, false, MODE, T>(d, t, in, U, out, shadow, partial, dtp, s, pull, wgs, Uold);
    } else {
        if (wgs) diffusion_tl<8, true, MODE, T>(d, t, in, U, out, shadow, partial, dtp, s, pull, wgs, Uold);
        else diffusion_tl<8, false, MODE, T>(d, t, in, U, out, shadow, partial, dtp, s, pull, wgs, Uold);
    }
}

template <typename T, typename TIN>
void launch_diffusion(const Dims &d, const DevTables &t, const TIN *in, const double *U, double *out, float *shadow,
                      double *partial, const double *dtp, hipStream_t s, const int32_t *pull, int tile, int wgs,
                      const double *Uold) {
    if (sizeof(T) == 4 && (U != out || !Uold || d.ni < 2))
        throw std::runtime_error("FP32 diffusion: updates the state in place, needs U^n and ni >= 2");
    diffusion_any<0, T>(d, t, in, U, out, shadow, partial, dtp, s, pull, tile, wgs, Uold);
}

template <typename T>
void launch_diffusion_only(const Dims &d, const DevTables &t, const double *U, double *R, hipStream_t s, int tile,
                           int wgs) {
    diffusion_any<1, T>(d, t, U, U, R, (float *)nullptr, (double *)nullptr, (const double *)nullptr, s,
                        (const int32_t *)nullptr, tile, wgs, (const double *)nullptr);
}

template <typename TS>
static void exchange_with(const Dims &d, const DevGhosts &g, const int2 *copy, int n_copy, TS *state, hipStream_t s) {
    if (n_copy > 0)
        hipLaunchKernelGGL((k_exchange_copy<TS>), dim3((unsigned)((n_copy + 255) / 256)), dim3(256), 0, s, d, copy,
                           n_copy, state);
    if (g.n_lerp > 0)
        hipLaunchKernelGGL((k_exchange_lerp<TS>), dim3((unsigned)((g.n_lerp + 255) / 256)), dim3(256), 0, s, d, g, state);
    if (g.n_total > g.first)
        hipLaunchKernelGGL((k_exchange<TS, sizeof(TS) == 8>), dim3((unsigned)((g.n_total - g.first + 255) / 256)),
                           dim3(256), 0, s, d, g, state);
}

template <typename TS>
void launch_exchange(const Dims &d, const DevGhosts &g, TS *state, hipStream_t s) {
    exchange_with(d, g, g.copy, g.n_copy, state, s);
}

template <typename TS>
void launch_exchange_rest(const Dims &d, const DevGhosts &g, TS *state, hipStream_t s) {
    if (!g.pull) throw std::runtime_error("no pull table: use the full exchange");
    exchange_with(d, g, g.copy_rest, g.n_copy_rest, state, s);
}

template <typename TS>
void launch_boundary(const Dims &d, const DevBoundary &b, TS *state, hipStream_t s, const int32_t *pull) {
    if (b.n == 0) return;
    const long long n = (long long)b.n * d.ng * d.NJ * d.NK;
    hipLaunchKernelGGL((k_boundary<TS>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, b, state, pull);
}

template <typename TS>
void launch_copy_radial(const Dims &d, const DevBoundary &b, const TS *src, TS *dst, hipStream_t s) {
    if (b.n == 0) return;
    const long long n = (long long)b.n * d.ng * d.NJ * d.NK;
    hipLaunchKernelGGL((k_copy_radial<TS>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, b, src, dst);
}

// the partial minima are laid out as the diffusion with the same tile edge
// leaves them, one per workgroup
void launch_dt_partial(const Dims &d, const DevTables &t, const double *U, double *partial, hipStream_t s, int tile) {
    check_tile(tile);
    if (tile == 16) hipLaunchKernelGGL(k_dt_tile<16>, dim3(tiles(d, 16)), dim3(256), 0, s, d, t, U, partial);
    else hipLaunchKernelGGL(k_dt_tile<8>, dim3(tiles(d, 8)), dim3(64), 0, s, d, t, U, partial);
}

void launch_dt_final(const Dims &d, const double *partial, double *dtp, double cfl, hipStream_t s, int tile) {
    check_tile(tile);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, s, partial, (int)tiles(d, tile), cfl, dtp);
}

void launch_set_scalar(double *p, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, s, p, v);
}

void launch_state_diff(const Dims &d, const double *a, const double *b, double tol, unsigned long long *w, int pass,
                       hipStream_t s) {
    const dim3 grid((unsigned)d.nblocks * 8);
    if (pass == 0) hipLaunchKernelGGL(k_state_diff<0>, grid, dim3(256), 0, s, d, a, b, tol, w);
    else hipLaunchKernelGGL(k_state_diff<1>, grid, dim3(256), 0, s, d, a, b, tol, w);
}

template <typename TS>
void launch_nan_scan(const Dims &d, const TS *st, unsigned long long *w, hipStream_t s) {
    hipLaunchKernelGGL(k_nan_scan<TS>, dim3((unsigned)d.nblocks * 8), dim3(256), 0, s, d, st, w);
}
void launch_nan_scan_psi(const Dims &d, const double *psi, unsigned long long *w, hipStream_t s) {
    hipLaunchKernelGGL(k_nan_scan_psi, dim3((unsigned)d.nblocks), dim3(256), 0, s, d, psi, w);
}
template void launch_nan_scan<double>(const Dims &, const double *, unsigned long long *, hipStream_t);
template void launch_nan_scan<float>(const Dims &, const float *, unsigned long long *, hipStream_t);

void launch_to_f32(const double *src, float *dst, long long n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_f32, dim3(4096), dim3(256), 0, s, src, dst, n);
}

// ---------------------------------------------------------------- initial state

// Seeded synthetic initial state on the interior of every block (the
// configured runs' initial condition; no reference counterpart beyond
// 3_6/ModInitiation.f90:14-188, which sets other fields): rho1 = v = 0;
// s1 = s_amp * c_p * U(-1,1) with U from a splitmix64 hash of (seed, block,
// cell); B = curl A of a smooth random vector potential, the sum over modes
// of (k x a) cos(k.x + phase) evaluated in global Cartesian coordinates (Yang
// positions and components through the involution (x, y, z) -> (-x, z, y) of
// ModYinYang_CoordConv, 3_6/ModYinYang.f90:8-17), before its scaling.
// rows per block (3 x NI): c_p, sqrt(rho0), c_s.  wgmax[workgroup]: the
// largest |B| / (sqrt(rho0) c_s) of its cells, for the scale.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void __launch_bounds__(256) k_synthetic(Dims d, const double *__restrict__ rng,
                                                   const int32_t *__restrict__ yin, const double *__restrict__ rows,
                                                   const SynthMode *__restrict__ modes, int nm, unsigned long long seed,
                                                   double s_amp, double *__restrict__ U, double *__restrict__ wgmax) {
    __shared__ double red[4];
    const long long n3 = (long long)d.ni * d.nj * d.nk;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double ratio = 0.0;
    if (t < (long long)d.nblocks * n3) {
        const int b = (int)(t / n3);
        const long long c = t - (long long)b * n3;
        // Fortran indices, i fastest
        const int i = (int)(c % d.ni) + 1, j = (int)((c / d.ni) % d.nj) + 1, k = (int)(c / ((long long)d.ni * d.nj)) + 1;
        const double *rg = rng + 6 * (long long)b;
        auto centre = [](double lo, double hi, int n, int f) {
            return (lo * (double)((float)(n - f) + 0.5f) + hi * (double)((float)f - 0.5f)) / (double)n;
        };
        const double r = centre(rg[0], rg[1], d.ni, i), th = centre(rg[2], rg[3], d.nj, j);
        const double ph = centre(rg[4], rg[5], d.nk, k);
        const double sth = sin(th), cth = cos(th), sph = sin(ph), cph = cos(ph);
        const bool y = yin[b] != 0;
        const double p0 = r * sth * cph, p1 = r * sth * sph, p2 = r * cth;
        const double x0 = y ? p0 : -p0, x1 = y ? p1 : p2, x2 = y ? p2 : p1;
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        for (int m = 0; m < nm; m++) {
            const SynthMode &M = modes[m];
            const double cc = cos(M.k[0] * x0 + M.k[1] * x1 + M.k[2] * x2 + M.ph);
            g0 += M.a[0] * cc; g1 += M.a[1] * cc; g2 += M.a[2] * cc;
        }
        const double l0 = y ? g0 : -g0, l1 = y ? g1 : g2, l2 = y ? g2 : g1;
        const double Br = l0 * (sth * cph) + l1 * (sth * sph) + l2 * cth;
        const double Bt = l0 * (cth * cph) + l1 * (cth * sph) + l2 * (-sth);
        const double Bp = l0 * (-sph) + l1 * cph + l2 * 0.0;
        const int I = i + d.ng - 1;
        const double *rw = rows + 3 * (long long)b * d.NI;
        const double cp = rw[I], srho = rw[d.NI + I], cs = rw[2 * d.NI + I];
        ratio = sqrt(Br * Br + Bt * Bt + Bp * Bp) / srho / cs;
        const unsigned long long id = ((unsigned long long)b << 32) ^ (unsigned long long)((i * 131 + j) * 1031 + k);
        const double u = (double)(splitmix64(seed ^ (id * 0x2545F4914F6CDD1Dull)) >> 11) * (1.0 / 9007199254740992.0);
        const long long q = (long long)b * 8 * d.P + cell_off(d, I, j + d.ng - 1, k + d.ng - 1);
        U[q + 4 * d.P] = Br;
        U[q + 5 * d.P] = Bt;
        U[q + 6 * d.P] = Bp;
        U[q + 7 * d.P] = s_amp * cp * (2.0 * u - 1.0);
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(ratio, off, 64);
        ratio = o > ratio ? o : ratio;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ratio;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) mx = red[w] > mx ? red[w] : mx;
        wgmax[blockIdx.x] = mx;
    }
}

// scale = b_ratio / max over the workgroup maxima (0 where the field is 0),
// then B *= scale on every interior cell
__global__ void __launch_bounds__(256) k_synthetic_scale(Dims d, const double *__restrict__ wgmax, int nwg,
                                                         double b_ratio, double *__restrict__ U) {
    __shared__ double sc;
    if (threadIdx.x == 0) {
        double mx = 0.0;
        for (int w = 0; w < nwg; w++) mx = wgmax[w] > mx ? wgmax[w] : mx;
        sc = mx > 0.0 ? b_ratio / mx : 0.0;
    }
    __syncthreads();
    const double scale = sc;
    const long long n3 = (long long)d.ni * d.nj * d.nk;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < (long long)d.nblocks * n3;
         t += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(t / n3);
        const long long c = t - (long long)b * n3;
        const int i = (int)(c % d.ni), j = (int)((c / d.ni) % d.nj), k = (int)(c / ((long long)d.ni * d.nj));
        const long long q = (long long)b * 8 * d.P + cell_off(d, i + d.ng, j + d.ng, k + d.ng);
        for (int v = 4; v < 7; v++) U[q + v * d.P] *= scale;
    }
}

void launch_synthetic(const Dims &d, const double *rng, const int32_t *yin, const double *rows, const SynthMode *modes,
                      int nm, unsigned long long seed, double s_amp, double b_ratio, double *U, double *wgmax,
                      hipStream_t s) {
    const long long n = (long long)d.nblocks * d.ni * d.nj * d.nk;
    const int nwg = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_synthetic, dim3((unsigned)nwg), dim3(256), 0, s, d, rng, yin, rows, modes, nm, seed, s_amp, U,
                       wgmax);
    hipLaunchKernelGGL(k_synthetic_scale, dim3(2048), dim3(256), 0, s, d, wgmax, nwg, b_ratio, U);
}

// ---------------------------------------------------------------- output

// Corner values of ModSave_AllCells (3_6/ModSavePlot.f90:585-640) for blocks
// [b0, b0 + nb): one thread per corner (i, j, k), i fastest, writing the
// line's 12 numbers to out[(b - b0)][corner][12]: the corner's Cartesian
// position in solar radii, then the 9 variables of the reference's state
// averaged over the 8 surrounding cells (ghosts included, summed in the
// reference's order), with v and B turned to Cartesian components; psi is 0
// (the state holds no psi: no term writes it with DivB_option = 0).
// geo per block: face radii xi_F (ni+1), then sin and cos of the theta faces
// (nj+1 each) and of the phi faces (nk+1 each), from the host's libm.
__global__ void __launch_bounds__(256) k_allcells(Dims d, const double *__restrict__ U,
                                                  const double *__restrict__ geo, const int32_t *__restrict__ yin,
                                                  int b0, int nb, double rsun, double *__restrict__ out) {
    const int ci = d.ni + 1, cj = d.nj + 1, ck = d.nk + 1;
    const long long nc = (long long)ci * cj * ck;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nb * nc) return;
    const int lb = (int)(t / nc), b = b0 + lb;
    const long long c = t - (long long)lb * nc;
    const int i = (int)(c % ci) + 1, j = (int)((c / ci) % cj) + 1, k = (int)(c / ((long long)ci * cj)) + 1;
    const long long G = ci + 2 * cj + 2 * ck;
    const double *g = geo + (long long)b * G;
    const double r = g[i - 1];
    const double sth = g[ci + j - 1], cth = g[ci + cj + j - 1];
    const double sph = g[ci + 2 * cj + k - 1], cph = g[ci + 2 * cj + ck + k - 1];
    // padded index of Fortran index f: f + ng - 1
    const int I = i + d.ng - 1, J = j + d.ng - 1, K = k + d.ng - 1;
    auto at = [&](int v, int II, int JJ, int KK) -> double {
        return U[(long long)b * 8 * d.P + (long long)v * d.P + cell_off(d, II, JJ, KK)];
    };
    double o[12];
    const bool y = yin[b] != 0;
    if (y) {
        o[0] = (r * sth) * cph; o[1] = (r * sth) * sph; o[2] = r * cth;
    } else {
        o[0] = -((r * sth) * cph); o[1] = r * cth; o[2] = (r * sth) * sph;
    }
    o[0] = o[0] / rsun; o[1] = o[1] / rsun; o[2] = o[2] / rsun;
    #pragma unroll
    for (int v = 0; v < 8; v++) {
        double s = at(v, I - 1, J - 1, K - 1) + at(v, I, J - 1, K - 1);
        s = s + at(v, I - 1, J, K - 1);
        s = s + at(v, I, J, K - 1);
        s = s + at(v, I - 1, J - 1, K);
        s = s + at(v, I, J - 1, K);
        s = s + at(v, I - 1, J, K);
        s = s + at(v, I, J, K);
        o[3 + v] = 0.125 * s;
    }
    o[11] = 0.0;
    // (vr, vt, vp) and (br, bt, bp) to Cartesian (:614-637); Yang: (-x, z, y)
    #pragma unroll
    for (int q = 0; q < 2; q++) {
        double *w = o + 4 + 3 * q;
        const double x0 = ((w[0] * sth) * cph + (w[1] * cth) * cph) - w[2] * sph;
        const double y0 = ((w[0] * sth) * sph + (w[1] * cth) * sph) + w[2] * cph;
        const double z0 = w[0] * cth - w[1] * sth;
        if (y) { w[0] = x0; w[1] = y0; w[2] = z0; }
        else { w[0] = -x0; w[1] = z0; w[2] = y0; }
    }
    double *dst = out + t * 12;
    #pragma unroll
    for (int q = 0; q < 12; q++) dst[q] = o[q];
}

void launch_allcells(const Dims &d, const double *U, const double *geo, const int32_t *yin, int b0, int nb,
                     double rsun, double *out, hipStream_t s) {
    const long long n = (long long)nb * (d.ni + 1) * (d.nj + 1) * (d.nk + 1);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_allcells, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, U, geo, yin, b0, nb, rsun,
                       out);
}

// ---------------------------------------------------------------- GLM divergence cleaning


// interior cell c of all blocks (phi fastest): block, interior and padded indices
struct GlmCell {
    int b, i, j, k, I, J, K;
};
__device__ __forceinline__ GlmCell glm_cell(const Dims &d, long long c) {
    const long long nn = (long long)d.ni * d.nj * d.nk;
    GlmCell g;
    g.b = (int)(c / nn);
    long long r = c - (long long)g.b * nn;
    g.k = (int)(r % d.nk);
    r /= d.nk;
    g.j = (int)(r % d.nj);
    g.i = (int)(r / d.nj);
    g.I = g.i + d.ng; g.J = g.j + d.ng; g.K = g.k + d.ng;
    return g;
}

// ModDivB_GLM (3_6/ModDivB.f90:13-38) at one RK stage, one lane per interior
// cell, on the stage input `in` (radial boundaries applied) and psi_in:
//   div B    ModSpherical_div (3_6/ModSpherical.f90:104-159): A_r r^2 along r
//            times 1/r^2, plus (A_t sin(t) along theta + A_p along phi) times
//            1/sin(t), then 1/r;
//   grad psi ModSpherical_Grad_f (:201-239);
//   R(psi) = -c_h^2 div B - (c_h / h_LLL) psi, c_h = v_wave of the step's
//            state (ModWaveSpeed_Dynamo at the step's start), from Ch.
// psi_out = psi_n + dt R(psi) / div (3_6/ModAdvance.f90:88-93), and grad psi
// goes to G ([block][3][interior cells]) for the stage kernel to subtract
// from R(B).  psi_out may be psi_n (the fourth stage): each lane reads psi_n
// at its own cell only.
__global__ void __launch_bounds__(256) k_glm_stage(Dims d, DevTables t, const double *__restrict__ in,
                                                   const double *__restrict__ Ch, const double *__restrict__ psi_in,
                                                   const double *psi_n, double *psi_out, double *__restrict__ G,
                                                   const double *__restrict__ dtp, double div) {
    const long long nn = (long long)d.ni * d.nj * d.nk;
    const long long c = xcd_lane();   // r/theta neighbour planes in one XCD's L2
    if (c >= nn * d.nblocks) return;
    const GlmCell g = glm_cell(d, c);
    const int I = g.I, J = g.J, K = g.K;
    const BlockMeta mt = t.meta[g.b];
    const long long rb = (long long)g.b * d.NI, tb = (long long)g.b * d.NJ;
    const double *S = in + (long long)g.b * 8 * d.P;
    const double *Ps = psi_in + (long long)g.b * d.P;
    // K is interior: the cell and its r and theta neighbours are in the main
    // part (offsets q +- SI, q +- SJ); only the phi neighbours may be ghosts
    const long long q = cell_off(d, I, J, K), SI = d.SI, SJ = d.SJ;
    const long long km2 = cell_off(d, I, J, K - 2), km1 = cell_off(d, I, J, K - 1), kp1 = cell_off(d, I, J, K + 1),
                    kp2 = cell_off(d, I, J, K + 2);
    const double *Br = S + 4 * d.P, *Bt = S + 5 * d.P, *Bp = S + 6 * d.P;
    auto a1 = [&](int di) -> double {
        const double x = t.xi[rb + I + di];
        return Br[q + di * SI] * (x * x);
    };
    auto a2 = [&](int dj) -> double { return Bt[q + dj * SJ] * t.sint[tb + J + dj]; };
    const double rinv = t.rinv[rb + I], sinv = t.sinv[tb + J];
    const double dr = d4(a1(-2), a1(-1), a1(1), a1(2), mt.cr) * t.r2inv[rb + I];
    double d23 = d4(a2(-2), a2(-1), a2(1), a2(2), mt.ct);
    d23 = d23 + d4(Bp[km2], Bp[km1], Bp[kp1], Bp[kp2], mt.cp);
    d23 = d23 * sinv;
    d23 = d23 * rinv;
    const double divb = dr + d23;
    const double g1 = d4(Ps[q - 2 * SI], Ps[q - SI], Ps[q + SI], Ps[q + 2 * SI], mt.cr);
    const double g2 = d4(Ps[q - 2 * SJ], Ps[q - SJ], Ps[q + SJ], Ps[q + 2 * SJ], mt.ct) * rinv;
    double g3 = d4(Ps[km2], Ps[km1], Ps[kp1], Ps[kp2], mt.cp) * rinv;
    g3 = g3 * sinv;
    // c_h and c_h / h_LLL of the step (k_glm_speed)
    const long long gb = (long long)g.b * 3 * nn + ((long long)g.i * d.nj + g.j) * d.nk + g.k;
    const long long cb = (long long)g.b * 2 * nn + ((long long)g.i * d.nj + g.j) * d.nk + g.k;
    const double w = Ch[cb], wh = Ch[cb + nn];
    const double rpsi = -((w * w) * divb) - wh * Ps[q];
    const long long pq = (long long)g.b * d.P + q;
    psi_out[pq] = psi_n[pq] + (*dtp * rpsi) / div;
    G[gb] = g1;
    G[gb + nn] = g2;
    G[gb + 2 * nn] = g3;
}

// c_h = v_wave and c_h / h_LLL of every interior cell of the step's state U
// (ModWaveSpeed_Dynamo, 3_6/ModWaveSpeed.f90:8-27, and h_LLL,
// 3_6/ModBlock.f90:420-422, as k_dt_tile forms them), once per step for the
// four GLM stages: Ch[block][2][interior cells].
__global__ void __launch_bounds__(256) k_glm_speed(Dims d, DevTables t, const double *__restrict__ U,
                                                   double *__restrict__ Ch) {
    const long long nn = (long long)d.ni * d.nj * d.nk;
    const long long c = xcd_lane();   // r/theta neighbour planes in one XCD's L2
    if (c >= nn * d.nblocks) return;
    const GlmCell g = glm_cell(d, c);
    const int I = g.I, J = g.J, K = g.K;
    const BlockMeta mt = t.meta[g.b];
    const long long rb = (long long)g.b * d.NI, tb = (long long)g.b * d.NJ;
    const double *Ub = U + (long long)g.b * 8 * d.P + cell_off(d, I, J, K);
    const double vr = Ub[d.P], vt = Ub[2 * d.P], vp = Ub[3 * d.P];
    const double br = Ub[4 * d.P], bt = Ub[5 * d.P], bp = Ub[6 * d.P];
    double w = t.cs[rb + I] + sqrt(vr * vr + vt * vt + vp * vp);
    w = w + sqrt(br * br + bt * bt + bp * bp) / t.sqrt_rho0[rb + I];
    const double xi = t.xi[rb + I];
    double h = mt.dxi, h2 = xi * mt.dxj, h3 = (xi * t.sint[tb + J]) * mt.dxk;
    h = h < h2 ? h : h2;
    h = h < h3 ? h : h3;
    const long long cb = (long long)g.b * 2 * nn + ((long long)g.i * d.nj + g.j) * d.nk + g.k;
    Ch[cb] = w;
    Ch[cb + nn] = w / h;
}

// The artificial diffusion of psi after the fourth stage
// (3_6/ModDiffusion.f90:9-150, h = 2, the ninth variable; 3_6/ModAdvance.f90:
// 101-108): the faces of k_diffusion's FP64 instantiation, the speeds of the
// state U where the reference reads them (transverse index - ng), and
// psi_new = psi + (0.5 dt) R.
__global__ void __launch_bounds__(256) k_glm_diffusion(Dims d, DevTables t, const double *__restrict__ U,
                                                       const double *__restrict__ psi, double *__restrict__ psi_new,
                                                       const double *__restrict__ dtp) {
    const long long nn = (long long)d.ni * d.nj * d.nk;
    const long long c = xcd_lane();   // r/theta neighbour planes in one XCD's L2
    if (c >= nn * d.nblocks) return;
    const GlmCell g = glm_cell(d, c);
    const int I = g.I, J = g.J, K = g.K;
    const BlockMeta mt = t.meta[g.b];
    const long long rb = (long long)g.b * d.NI, tb = (long long)g.b * d.NJ;
    const double *Sb = U + (long long)g.b * 8 * d.P;
    const double *Pb = psi + (long long)g.b * d.P;
    const double *cs = t.cs + rb;
    // the cell and its r and theta neighbours are in the main part; the phi
    // neighbours and the speed cells at K - 2 may be ghosts
    const long long q = cell_off(d, I, J, K), SI = d.SI, SJ = d.SJ;
    const long long km2 = cell_off(d, I, J, K - 2), km1 = cell_off(d, I, J, K - 1), kp1 = cell_off(d, I, J, K + 1),
                    kp2 = cell_off(d, I, J, K + 2);
    const double *V1 = Sb + d.P, *V2 = Sb + 2 * d.P, *V3 = Sb + 3 * d.P;
    // speed cells (transverse index - ng): (I', J - 2, K - 2), (I - 2, J', K - 2),
    // (I - 2, J - 2, K'); K - 2 lies in the side part below the interior, whose
    // plane and row pitches are NJ 2ng and 2ng
    const long long s1 = cell_off(d, I, J - 2, K - 2), s2 = cell_off(d, I - 2, J, K - 2);
    const bool kmain = K - 2 >= d.ng;
    const long long pI = kmain ? SI : (long long)d.NJ * 2 * d.ng, pJ = kmain ? SJ : 2 * d.ng;
    const double c1m = fabs(V1[s1 - pI]) + cs[I - 1];
    const double c10 = fabs(V1[s1]) + cs[I];
    const double c1p = fabs(V1[s1 + pI]) + cs[I + 1];
    const double p0 = Pb[q];
    const double f0 = (face_flux(Pb[q - 2 * SI], Pb[q - SI], p0, Pb[q + SI]) * (c10 + c1m)) * 0.5;
    const double f1 = (face_flux(Pb[q - SI], p0, Pb[q + SI], Pb[q + 2 * SI]) * (c1p + c10)) * 0.5;
    const double csI2 = cs[I - 2];
    const double c2m = fabs(V2[s2 - pJ]) + csI2, c20 = fabs(V2[s2]) + csI2, c2p = fabs(V2[s2 + pJ]) + csI2;
    const double fj0 = (face_flux(Pb[q - 2 * SJ], Pb[q - SJ], p0, Pb[q + SJ]) * (c20 + c2m)) * 0.5;
    const double fj1 = (face_flux(Pb[q - SJ], p0, Pb[q + SJ], Pb[q + 2 * SJ]) * (c2p + c20)) * 0.5;
    const double c3m = fabs(V3[cell_off(d, I - 2, J - 2, K - 1)]) + csI2, c30 = fabs(V3[cell_off(d, I - 2, J - 2, K)]) + csI2,
                 c3p = fabs(V3[cell_off(d, I - 2, J - 2, K + 1)]) + csI2;
    const double fk0 = (face_flux(Pb[km2], Pb[km1], p0, Pb[kp1]) * (c30 + c3m)) * 0.5;
    const double fk1 = (face_flux(Pb[km1], p0, Pb[kp1], Pb[kp2]) * (c3p + c30)) * 0.5;
    const double xic = t.xi[rb + I];
    const double L2 = mt.dxj * xic;
    const double L3 = (mt.dxk * xic) * t.sint[tb + J];
    double r = 0.0;
    r = r + (f0 - f1) / mt.dxi;
    r = r + (fj0 - fj1) / L2;
    r = r + (fk0 - fk1) / L3;
    psi_new[(long long)g.b * d.P + q] = p0 + (0.5 * *dtp) * r;
}

static unsigned glm_grid(const Dims &d) {
    const long long n = (long long)d.ni * d.nj * d.nk * d.nblocks;
    return (unsigned)((n + 255) / 256);
}

void launch_glm_speed(const Dims &d, const DevTables &t, const double *U, double *Ch, hipStream_t s) {
    hipLaunchKernelGGL(k_glm_speed, dim3(glm_grid(d)), dim3(256), 0, s, d, t, U, Ch);
}

void launch_glm_stage(const Dims &d, const DevTables &t, const double *in, const double *Ch, const double *psi_in,
                      const double *psi_n, double *psi_out, double *G, const double *dtp, double div, hipStream_t s) {
    if (psi_in == psi_out) throw std::runtime_error("GLM stage: psi_in and psi_out must differ");
    hipLaunchKernelGGL(k_glm_stage, dim3(glm_grid(d)), dim3(256), 0, s, d, t, in, Ch, psi_in, psi_n, psi_out, G, dtp,
                       div);
}

void launch_glm_diffusion(const Dims &d, const DevTables &t, const double *U, const double *psi, double *psi_new,
                          const double *dtp, hipStream_t s) {
    if (psi == psi_new) throw std::runtime_error("GLM diffusion: psi and psi_new must differ");
    hipLaunchKernelGGL(k_glm_diffusion, dim3(glm_grid(d)), dim3(256), 0, s, d, t, U, psi, psi_new, dtp);
}

void launch_exchange_glm(const Dims &d, const DevGhosts &g, double *state, double *psi, hipStream_t s) {
    if (g.n_lerp > 0) throw std::runtime_error("GLM ghost fill: the FP64 ghost tables have no r-only entries");
    if (g.n_copy > 0)
        hipLaunchKernelGGL((k_exchange_copy<double, true>), dim3((unsigned)((g.n_copy + 255) / 256)), dim3(256), 0, s, d,
                           g.copy, g.n_copy, state, psi);
    if (g.n_total > g.first)
        hipLaunchKernelGGL((k_exchange<double, true, true>), dim3((unsigned)((g.n_total - g.first + 255) / 256)),
                           dim3(256), 0, s, d, g, state, psi);
}

// explicit instantiations: FP64 path, FP32-deviation path
template void launch_stage<double, double, double, double>(const Dims &, const DevTables &, const double *, const double *, double *, float *, const double *, double, hipStream_t, const int32_t *, int, int, const double *);
template void launch_stage<float, float, float, float>(const Dims &, const DevTables &, const float *, const float *, float *, float *, const double *, double, hipStream_t, const int32_t *, int, int, const double *);
template void launch_stage<float, float, double, double>(const Dims &, const DevTables &, const float *, const double *, double *, float *, const double *, double, hipStream_t, const int32_t *, int, int, const double *);
template void launch_rhs_only<double, double>(const Dims &, const DevTables &, const double *, double *, hipStream_t, int, int);
template void launch_rhs_only<float, double>(const Dims &, const DevTables &, const double *, double *, hipStream_t, int, int);
template void launch_diffusion<double, double>(const Dims &, const DevTables &, const double *, const double *, double *, float *, double *, const double *, hipStream_t, const int32_t *, int, int, const double *);
template void launch_diffusion<float, float>(const Dims &, const DevTables &, const float *, const double *, double *, float *, double *, const double *, hipStream_t, const int32_t *, int, int, const double *);
template void launch_diffusion_only<double>(const Dims &, const DevTables &, const double *, double *, hipStream_t, int, int);
template void launch_diffusion_only<float>(const Dims &, const DevTables &, const double *, double *, hipStream_t, int, int);
template void launch_exchange<double>(const Dims &, const DevGhosts &, double *, hipStream_t);
template void launch_exchange<float>(const Dims &, const DevGhosts &, float *, hipStream_t);
template void launch_exchange_rest<float>(const Dims &, const DevGhosts &, float *, hipStream_t);
template void launch_boundary<double>(const Dims &, const DevBoundary &, double *, hipStream_t, const int32_t *);
template void launch_boundary<float>(const Dims &, const DevBoundary &, float *, hipStream_t, const int32_t *);
template void launch_copy_radial<double>(const Dims &, const DevBoundary &, const double *, double *, hipStream_t);
template void launch_copy_radial<float>(const Dims &, const DevBoundary &, const float *, float *, hipStream_t);

}  // namespace falcon
