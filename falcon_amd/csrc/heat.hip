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
// here table-driven), summed restriction of the residual, injection
// prolongation, damped Jacobi (ModMultigrid_Jacobi_Mz_r, 3_5/ModMultigrid.f90:304-325,
// with damping).  Krylov: flexible CG (FCG(1)), everything FP64; dot products
// are deterministic two-pass reductions (wavefront shuffle, LDS, second pass).
//
// All kernels carry a system index in blockIdx.y: nsys independent systems
// (same grid, own dt and right-hand side) are solved in the same launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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
    int m, nb;
    long long P;            // padded cells per block
    long long M3;           // interior cells per block
    const double *V, *Ai, *Aj, *Ak, *off;
    const int32_t *gdst, *gsrc;
    const double *gw;
    int ng;                 // ghost entries
};

__device__ __forceinline__ long long pad_index(int m, int i, int j, int k) {
    return ((long long)(i + 1) * (m + 2) + (j + 1)) * (m + 2) + (k + 1);
}

// face-ghost fill of one level: trilinear inter/extrapolation inside the donor
// block, term order of ModMath.f90:283-291 (weights may lie in [-0.5, 1.5])
__global__ void __launch_bounds__(256) k_hx(HLev L, double *__restrict__ x, long long sys_stride) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.ng) return;
    double *X = x + (long long)blockIdx.y * sys_stride;
    const int S1 = L.m + 2, SI = S1 * S1;
    const int src = L.gsrc[e];
    const double wi = L.gw[3 * e], wj = L.gw[3 * e + 1], wk = L.gw[3 * e + 2];
    const double *F = X + src;
    double t = F[0] * (1.0 - wi) * (1.0 - wj) * (1.0 - wk);
    t = t + F[SI] * (wi) * (1.0 - wj) * (1.0 - wk);
    t = t + F[S1] * (1.0 - wi) * (wj) * (1.0 - wk);
    t = t + F[1] * (1.0 - wi) * (1.0 - wj) * (wk);
    t = t + F[SI + S1] * (wi) * (wj) * (1.0 - wk);
    t = t + F[S1 + 1] * (1.0 - wi) * (wj) * (wk);
    t = t + F[SI + 1] * (wi) * (1.0 - wj) * (wk);
    t = t + F[SI + S1 + 1] * (wi) * (wj) * (wk);
    X[L.gdst[e]] = t;
}

// A x on the interior.  MODE 0: y = A x;  MODE 1: y = r - A x;
// MODE 2: damped Jacobi, y = x + omega (r - A x) / diag (padded output).
template <int MODE>
__global__ void __launch_bounds__(256) k_happly(HLev L, const double *__restrict__ x, const double *__restrict__ r,
                                                double *__restrict__ y, const double *__restrict__ inv_dt,
                                                long long pstride, long long istride, double omega) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)L.nb * L.M3) return;
    const int s = blockIdx.y;
    const int m = L.m;
    const int b = (int)(c / L.M3);
    const int q = (int)(c - (long long)b * L.M3);
    const int i = q / (m * m), j = (q / m) % m, k = q % m;
    const double *X = x + s * pstride + (long long)b * L.P;
    const long long S1 = m + 2, SI = S1 * S1;
    const long long pc = pad_index(m, i, j, k);
    const double *ai = L.Ai + (long long)b * (m + 1) * m * m;
    const double *aj = L.Aj + (long long)b * m * (m + 1) * m;
    const double *ak = L.Ak + (long long)b * m * m * (m + 1);
    const double cim = ai[((long long)i * m + j) * m + k], cip = ai[((long long)(i + 1) * m + j) * m + k];
    const double cjm = aj[((long long)i * (m + 1) + j) * m + k], cjp = aj[((long long)i * (m + 1) + j + 1) * m + k];
    const double ckm = ak[((long long)i * m + j) * (m + 1) + k], ckp = ak[((long long)i * m + j) * (m + 1) + k + 1];
    const double diag = L.V[(long long)b * L.M3 + q] * inv_dt[s] + L.off[(long long)b * L.M3 + q];
    const double xc = X[pc];
    double nb = cim * X[pc - SI] + cip * X[pc + SI];
    nb = nb + cjm * X[pc - S1] + cjp * X[pc + S1];
    nb = nb + ckm * X[pc - 1] + ckp * X[pc + 1];
    const double ax = diag * xc - nb;
    const long long io = s * istride + c;
    if (MODE == 0) y[io] = ax;
    else if (MODE == 1) y[io] = r[io] - ax;
    else y[s * pstride + (long long)b * L.P + pc] = xc + omega * (r[io] - ax) / diag;
}

// residual restriction: coarse cell = sum of its 8 children (integrated residual)
__global__ void __launch_bounds__(256) k_hrestrict(int mf, int nb, const double *__restrict__ rf, double *__restrict__ rc,
                                                   long long fstride, long long cstride) {
    const int mc = mf / 2;
    const long long Mc = (long long)mc * mc * mc, Mf = (long long)mf * mf * mf;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nb * Mc) return;
    const int s = blockIdx.y;
    const int b = (int)(c / Mc);
    const int q = (int)(c - b * Mc);
    const int i = q / (mc * mc), j = (q / mc) % mc, k = q % mc;
    const double *R = rf + s * fstride + b * Mf;
    double t = 0.0;
    for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++)
            for (int dk = 0; dk < 2; dk++) t += R[((long long)(2 * i + di) * mf + (2 * j + dj)) * mf + (2 * k + dk)];
    rc[s * cstride + c] = t;
}

// prolongation by injection: fine z += coarse z of the parent cell (padded arrays)
__global__ void __launch_bounds__(256) k_hprolong(int mf, int nb, const double *__restrict__ zc, double *__restrict__ zf,
                                                  long long fstride, long long cstride) {
    const int mc = mf / 2;
    const long long Mf = (long long)mf * mf * mf;
    const long long Pf = (long long)(mf + 2) * (mf + 2) * (mf + 2), Pc = (long long)(mc + 2) * (mc + 2) * (mc + 2);
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nb * Mf) return;
    const int s = blockIdx.y;
    const int b = (int)(c / Mf);
    const int q = (int)(c - b * Mf);
    const int i = q / (mf * mf), j = (q / mf) % mf, k = q % mf;
    zf[s * fstride + b * Pf + pad_index(mf, i, j, k)] += zc[s * cstride + b * Pc + pad_index(mc, i / 2, j / 2, k / 2)];
}

// interior <-> padded copies and fills
__global__ void __launch_bounds__(256) k_hzero(double *__restrict__ a, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// up to 3 dot products per system over the interior: a_p . b_p, with a or b padded
// (flags) or interior.  partial[s][p][wg]
struct DotArgs {
    const double *a[3], *b[3];
    int apad[3], bpad[3];
    int npairs;
};

__global__ void __launch_bounds__(256) k_hdot(HLev L, DotArgs D, double *__restrict__ partial, long long pstride,
                                              long long istride) {
    __shared__ double red[3][4];
    const int s = blockIdx.y;
    const int m = L.m;
    const long long N = (long long)L.nb * L.M3;
    double acc[3] = {0.0, 0.0, 0.0};
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(c / L.M3);
        const int q = (int)(c - (long long)b * L.M3);
        const long long pp = s * pstride + (long long)b * L.P + pad_index(m, q / (m * m), (q / m) % m, q % m);
        const long long ii = s * istride + c;
        for (int p = 0; p < D.npairs; p++) {
            const double x = D.apad[p] ? D.a[p][pp] : D.a[p][ii];
            const double y = D.bpad[p] ? D.b[p][pp] : D.b[p][ii];
            acc[p] += x * y;
        }
    }
    for (int p = 0; p < D.npairs; p++) {
        double v = acc[p];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0) red[p][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < D.npairs) {
        const int p = threadIdx.x;
        partial[((long long)s * 3 + p) * gridDim.x + blockIdx.x] = ((red[p][0] + red[p][1]) + red[p][2]) + red[p][3];
    }
}

__global__ void __launch_bounds__(64) k_hdot_final(const double *__restrict__ partial, int nwg, int npairs,
                                                   double *__restrict__ out) {
    const int s = blockIdx.x;
    for (int p = 0; p < npairs; p++) {
        double v = 0.0;
        for (int w = threadIdx.x; w < nwg; w += 64) v += partial[((long long)s * 3 + p) * nwg + w];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (threadIdx.x == 0) out[s * 3 + p] = v;
    }
}

// Per-system scalars on the device (one thread per system), kSc slots each:
// [0] alpha = <p,r>/<p,q>, [1] beta = <z,q>/gamma, [2] gamma = <p,q>,
// [3] feature scale of the learned correction, [4] its step along c.
constexpr int kSc = 8;
// what 0: alpha and gamma (dots: [0]=<p,q>, [1]=<p,r>)
// what 1: beta (dots: [0]=<z,q>)
// what 2: correction step <c,r'>/<c,Ac> (dots: [0]=<c,r'>, [1]=<c,Ac>), 0 unless <c,Ac> > 0
// what 3: feature scale sqrt(sum / ncell) (dots: [0]=sum (r'/diag)^2)
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
        S[3] = sqrt(dots[3 * s] / ncell);
    } else {
        S[5] = dots[3 * s];                 // <b, b>
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

// x += alpha p ; r -= alpha q (p padded, x, r, q interior)
__global__ void __launch_bounds__(256) k_hupdate_xr(HLev L, double *__restrict__ x, double *__restrict__ r,
                                                    const double *__restrict__ p, const double *__restrict__ q,
                                                    const double *__restrict__ sc, long long pstride, long long istride) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)L.nb * L.M3) return;
    const int s = blockIdx.y;
    const double a = sc[kSc * s];
    const int m = L.m;
    const int b = (int)(c / L.M3);
    const int qq = (int)(c - (long long)b * L.M3);
    const double pv = p[s * pstride + (long long)b * L.P + pad_index(m, qq / (m * m), (qq / m) % m, qq % m)];
    const long long io = s * istride + c;
    x[io] = x[io] + a * pv;
    r[io] = r[io] - a * q[io];
}

// p = z - beta p (padded interiors); beta = 0 starts a fresh direction
__global__ void __launch_bounds__(256) k_hupdate_p(HLev L, double *__restrict__ p, const double *__restrict__ z,
                                                   const double *__restrict__ sc, long long pstride, int use_beta) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)L.nb * L.M3) return;
    const int s = blockIdx.y;
    const double be = use_beta ? sc[kSc * s + 1] : 0.0;
    const int m = L.m;
    const int b = (int)(c / L.M3);
    const int qq = (int)(c - (long long)b * L.M3);
    const long long o = s * pstride + (long long)b * L.P + pad_index(m, qq / (m * m), (qq / m) % m, qq % m);
    p[o] = z[o] - be * p[o];
}

// b = (V/dt) T for the interior (T interior layout)
__global__ void __launch_bounds__(256) k_hrhs(HLev L, const double *__restrict__ T, double *__restrict__ bvec,
                                              const double *__restrict__ inv_dt, long long istride) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)L.nb * L.M3) return;
    const int s = blockIdx.y;
    bvec[s * istride + c] = L.V[c] * inv_dt[s] * T[s * istride + c];
}

__global__ void __launch_bounds__(256) k_hcopy(const double *__restrict__ a, double *__restrict__ b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ---- learned correction of the V-cycle
// Block partial sums of up to 3 per-thread values into partial[s][p][blockIdx.x].
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

__device__ __forceinline__ double apply_cell(const HLev &L, const double *__restrict__ X, int b, int q, double inv_dt,
                                             double &diag) {
    const int m = L.m;
    const int i = q / (m * m), j = (q / m) % m, k = q % m;
    const long long S1 = m + 2, SI = S1 * S1;
    const long long pc = pad_index(m, i, j, k);
    const double *ai = L.Ai + (long long)b * (m + 1) * m * m;
    const double *aj = L.Aj + (long long)b * m * (m + 1) * m;
    const double *ak = L.Ak + (long long)b * m * m * (m + 1);
    const double cim = ai[((long long)i * m + j) * m + k], cip = ai[((long long)(i + 1) * m + j) * m + k];
    const double cjm = aj[((long long)i * (m + 1) + j) * m + k], cjp = aj[((long long)i * (m + 1) + j + 1) * m + k];
    const double ckm = ak[((long long)i * m + j) * (m + 1) + k], ckp = ak[((long long)i * m + j) * (m + 1) + k + 1];
    diag = L.V[(long long)b * L.M3 + q] * inv_dt + L.off[(long long)b * L.M3 + q];
    double nb = cim * X[pc - SI] + cip * X[pc + SI];
    nb = nb + cjm * X[pc - S1] + cjp * X[pc + S1];
    nb = nb + ckm * X[pc - 1] + ckp * X[pc + 1];
    return diag * X[pc] - nb;
}

// Az = A z, r' = r - A z, network input u = r'/diag (float, unscaled) and
// partial sums of u^2 (z padded with fresh ghosts)
__global__ void __launch_bounds__(256) k_hfeat(HLev L, const double *__restrict__ z, const double *__restrict__ r,
                                               double *__restrict__ Az, double *__restrict__ rp, float *__restrict__ u,
                                               const double *__restrict__ inv_dt, double *__restrict__ partial,
                                               long long pstride, long long istride) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    double v[1] = {0.0};
    if (c < (long long)L.nb * L.M3) {
        const int b = (int)(c / L.M3);
        const int q = (int)(c - (long long)b * L.M3);
        double diag;
        const double ax = apply_cell(L, z + s * pstride + (long long)b * L.P, b, q, inv_dt[s], diag);
        const long long io = s * istride + c;
        const double res = r[io] - ax;
        Az[io] = ax;
        rp[io] = res;
        const double w = res / diag;
        u[io] = (float)w;
        v[0] = w * w;
    }
    block_partials<1>(v, partial, s);
}

// Ac = A c and partial sums of c.r' and c.Ac (c padded with fresh ghosts)
__global__ void __launch_bounds__(256) k_happly_dot(HLev L, const double *__restrict__ cpad, const double *__restrict__ rp,
                                                    double *__restrict__ Ac, const double *__restrict__ inv_dt,
                                                    double *__restrict__ partial, long long pstride, long long istride) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    double v[2] = {0.0, 0.0};
    if (c < (long long)L.nb * L.M3) {
        const int b = (int)(c / L.M3);
        const int q = (int)(c - (long long)b * L.M3);
        const double *C = cpad + s * pstride + (long long)b * L.P;
        double diag;
        const double ac = apply_cell(L, C, b, q, inv_dt[s], diag);
        const int m = L.m;
        const double cv = C[pad_index(m, q / (m * m), (q / m) % m, q % m)];
        const long long io = s * istride + c;
        Ac[io] = ac;
        v[0] = cv * rp[io];
        v[1] = cv * ac;
    }
    block_partials<2>(v, partial, s);
}

// z += a c (padded interiors), Az += a Ac, partial sums of z.q (q = A p of the
// previous direction), a = correction step of the system
__global__ void __launch_bounds__(256) k_hzcorr(HLev L, double *__restrict__ z, const double *__restrict__ cpad,
                                                double *__restrict__ Az, const double *__restrict__ Ac,
                                                const double *__restrict__ q, const double *__restrict__ sc,
                                                double *__restrict__ partial, long long pstride, long long istride) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    double v[1] = {0.0};
    if (c < (long long)L.nb * L.M3) {
        const double a = sc[kSc * s + 4];
        const int m = L.m;
        const int b = (int)(c / L.M3);
        const int qq = (int)(c - (long long)b * L.M3);
        const long long o = s * pstride + (long long)b * L.P + pad_index(m, qq / (m * m), (qq / m) % m, qq % m);
        const double zn = z[o] + a * cpad[o];
        z[o] = zn;
        const long long io = s * istride + c;
        Az[io] = Az[io] + a * Ac[io];
        v[0] = zn * q[io];
    }
    block_partials<1>(v, partial, s);
}

// p = z - beta p (padded interiors) and q = Az - beta q: the new direction and
// its image without an operator apply
__global__ void __launch_bounds__(256) k_hupdate_pq(HLev L, double *__restrict__ p, const double *__restrict__ z,
                                                    double *__restrict__ q, const double *__restrict__ Az,
                                                    const double *__restrict__ sc, long long pstride, long long istride,
                                                    int use_beta) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)L.nb * L.M3) return;
    const int s = blockIdx.y;
    const double be = use_beta ? sc[kSc * s + 1] : 0.0;
    const int m = L.m;
    const int b = (int)(c / L.M3);
    const int qq = (int)(c - (long long)b * L.M3);
    const long long o = s * pstride + (long long)b * L.P + pad_index(m, qq / (m * m), (qq / m) % m, qq % m);
    p[o] = z[o] - be * p[o];
    const long long io = s * istride + c;
    q[io] = Az[io] - be * q[io];
}

// training records: copy the network input image and w = x + z of record slot
// r (system rs[r], block rb[r]) into X[r] (float) and W[r] (double)
__global__ void __launch_bounds__(256) k_hrecord(HLev L, const float *__restrict__ xin, const double *__restrict__ x,
                                                 const double *__restrict__ z, const int *__restrict__ rs,
                                                 const int *__restrict__ rb, float *__restrict__ X, double *__restrict__ W,
                                                 const double *__restrict__ sc, double *__restrict__ rscale,
                                                 long long pstride, long long istride) {
    const long long r = blockIdx.x;
    const int s = rs[r], b = rb[r], m = L.m;
    if (threadIdx.x == 0) rscale[r] = sc[kSc * s + 3];
    for (int q = threadIdx.x; q < (int)L.M3; q += blockDim.x) {
        const long long io = s * istride + (long long)b * L.M3 + q;
        const double scale = sc[kSc * s + 3];
        X[r * L.M3 + q] = scale > 0.0 ? (float)((double)xin[io] / scale) : 0.0f;
        W[r * L.M3 + q] = x[io] + z[s * pstride + (long long)b * L.P + pad_index(m, q / (m * m), (q / m) % m, q % m)];
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
    // per level work arrays (device)
    std::vector<double *> Z, Z2, R, RES;
    // level-0 FCG arrays
    double *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *b = nullptr, *tmp = nullptr;
    double *inv_dt = nullptr, *partial = nullptr, *dots = nullptr, *sc = nullptr;
    int *active = nullptr;
    // learned correction work arrays (allocated on first use)
    double *Az = nullptr, *rp = nullptr, *Ac = nullptr, *cpad = nullptr, *partial_cells = nullptr;
    float *xin = nullptr;
    int learned_pre0 = 2, learned_post0 = 2;
    // captured FCG iteration per mode (plain, learned) and device-side convergence state
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    bool use_graph = true;
    int *iters_dev = nullptr, *nactive_dev = nullptr;
    double *relres_dev = nullptr;
    void drop_graphs() {
        for (auto &g : graph)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    }
    unsigned ncell_blocks() const { return (unsigned)((ist(0) + 255) / 256); }
    void ensure_learned();
    void features(double *z);
    void learned_correction(double *z);
    void final_dots(int npairs, unsigned nwg) {
        hipLaunchKernelGGL(k_hdot_final, dim3((unsigned)nsys), dim3(64), 0, stream, (const double *)partial_cells, (int)nwg,
                           npairs, dots);
    }
    int dot_wg = 1024;
    int pre = 2, post = 2, coarse_sweeps = 8;
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
    long long pst(int l) const { return (long long)nb * L[l].P; }
    long long ist(int l) const { return (long long)nb * L[l].M3; }
    dim3 grid_cells(int l) const { return dim3((unsigned)((nb * L[l].M3 + 255) / 256), (unsigned)nsys); }

    void exchange(int l, double *x_) {
        if (L[l].ng == 0) return;
        hipLaunchKernelGGL(k_hx, dim3((unsigned)((L[l].ng + 255) / 256), (unsigned)nsys), dim3(256), 0, stream, L[l], x_,
                           pst(l));
    }
    void zero(double *a, long long n) {
        hipLaunchKernelGGL(k_hzero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, n);
    }
    void jacobi(int l, double *&z, double *&z2, const double *rr) {
        exchange(l, z);
        hipLaunchKernelGGL((k_happly<2>), grid_cells(l), dim3(256), 0, stream, L[l], z, rr, z2, inv_dt_l(l), pst(l),
                           ist(l), omega);
        std::swap(z, z2);
    }
    // coarse levels see the same dt (the operator is rediscretised per level)
    const double *inv_dt_l(int) const { return inv_dt; }

    // V-cycle: z0 (padded, level 0) = M r0 (interior, level 0)
    // pre0 / post0: smoothing sweeps on the finest level (others use pre / post)
    void vcycle(const double *r0, double *&z0, int pre0 = -1, int post0 = -1) {
        if (pre0 < 0) pre0 = pre;
        if (post0 < 0) post0 = post;
        R[0] = const_cast<double *>(r0);
        Z[0] = z0;
        for (int l = 0; l < nlev; l++) {
            zero(Z[l], pst(l) * nsys);
            const int sweeps = (l == nlev - 1) ? coarse_sweeps : (l == 0 ? pre0 : pre);
            for (int s = 0; s < sweeps; s++) jacobi(l, Z[l], Z2[l], R[l]);
            if (l == nlev - 1) break;
            exchange(l, Z[l]);
            hipLaunchKernelGGL((k_happly<1>), grid_cells(l), dim3(256), 0, stream, L[l], Z[l], R[l], RES[l], inv_dt,
                               pst(l), ist(l), 0.0);
            hipLaunchKernelGGL(k_hrestrict, grid_cells(l + 1), dim3(256), 0, stream, L[l].m, nb, RES[l], R[l + 1], ist(l),
                               ist(l + 1));
        }
        for (int l = nlev - 2; l >= 0; l--) {
            hipLaunchKernelGGL(k_hprolong, grid_cells(l), dim3(256), 0, stream, L[l].m, nb, Z[l + 1], Z[l], pst(l),
                               pst(l + 1));
            const int sweeps = l == 0 ? post0 : post;
            for (int s = 0; s < sweeps; s++) jacobi(l, Z[l], Z2[l], R[l]);
        }
        z0 = Z[0];
    }

    void dot(int l, int npairs, const double *a0, int ap0, const double *b0, int bp0, const double *a1 = nullptr,
             int ap1 = 0, const double *b1 = nullptr, int bp1 = 0) {
        DotArgs D{};
        D.a[0] = a0; D.apad[0] = ap0; D.b[0] = b0; D.bpad[0] = bp0;
        D.a[1] = a1; D.apad[1] = ap1; D.b[1] = b1; D.bpad[1] = bp1;
        D.npairs = npairs;
        hipLaunchKernelGGL(k_hdot, dim3((unsigned)dot_wg, (unsigned)nsys), dim3(256), 0, stream, L[l], D, partial, pst(l),
                           ist(l));
        hipLaunchKernelGGL(k_hdot_final, dim3((unsigned)nsys), dim3(64), 0, stream, (const double *)partial, dot_wg, npairs,
                           dots);
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
                    const double kap = hp.kappa0 / (x * x * x);
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
        D.P = (long long)(H.m + 2) * (H.m + 2) * (H.m + 2);
        D.M3 = (long long)H.m * H.m * H.m;
        D.V = I.upload(H.V); D.Ai = I.upload(H.Ai); D.Aj = I.upload(H.Aj); D.Ak = I.upload(H.Ak);
        D.off = I.upload(H.offsum);
        D.gdst = I.upload(H.g_dst); D.gsrc = I.upload(H.g_src); D.gw = I.upload(H.g_w);
        D.ng = (int)H.g_dst.size();
    }
    I.Z.resize(I.nlev); I.Z2.resize(I.nlev); I.R.resize(I.nlev); I.RES.resize(I.nlev);
    for (int l = 0; l < I.nlev; l++) {
        I.Z[l] = I.dalloc(I.pst(l) * nsys);
        I.Z2[l] = I.dalloc(I.pst(l) * nsys);
        I.R[l] = l ? I.dalloc(I.ist(l) * nsys) : nullptr;
        I.RES[l] = I.dalloc(I.ist(l) * nsys);
    }
    I.x = I.dalloc(I.ist(0) * nsys);
    I.r = I.dalloc(I.ist(0) * nsys);
    I.q = I.dalloc(I.ist(0) * nsys);
    I.b = I.dalloc(I.ist(0) * nsys);
    I.tmp = I.dalloc(I.ist(0) * nsys);
    I.p = I.dalloc(I.pst(0) * nsys);
    I.inv_dt = I.dalloc(nsys);
    I.partial = I.dalloc((size_t)nsys * 3 * I.dot_wg);
    I.dots = I.dalloc((size_t)nsys * 3);
    I.sc = I.dalloc((size_t)nsys * kSc);
    void *a = nullptr;
    HCHECK(hipMalloc(&a, sizeof(int) * nsys));
    I.allocs.push_back(a);
    I.active = (int *)a;
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

void HeatSolver::apply(const std::vector<double> &x, std::vector<double> &y) {
    // y = A x for system-major interior vectors (parity tests)
    Impl &I = *d;
    const long long n = I.ist(0) * I.nsys;
    if ((long long)x.size() != n) throw std::runtime_error("x has the wrong size");
    std::vector<double> pad(I.pst(0) * I.nsys, 0.0);
    const int m = I.L[0].m;
    for (int s = 0; s < I.nsys; s++)
        for (int bb = 0; bb < I.nb; bb++)
            for (int c = 0; c < m * m * m; c++) {
                const int i = c / (m * m), j = (c / m) % m, k = c % m;
                pad[s * I.pst(0) + (long long)bb * I.L[0].P + ((long long)(i + 1) * (m + 2) + (j + 1)) * (m + 2) + k + 1] =
                    x[s * I.ist(0) + (long long)bb * m * m * m + c];
            }
    HCHECK(hipMemcpy(I.p, pad.data(), sizeof(double) * pad.size(), hipMemcpyHostToDevice));
    I.exchange(0, I.p);
    hipLaunchKernelGGL((k_happly<0>), I.grid_cells(0), dim3(256), 0, I.stream, I.L[0], (const double *)I.p,
                       (const double *)nullptr, I.q, (const double *)I.inv_dt, I.pst(0), I.ist(0), 0.0);
    y.resize(n);
    HCHECK(hipStreamSynchronize(I.stream));
    HCHECK(hipMemcpy(y.data(), I.q, sizeof(double) * n, hipMemcpyDeviceToHost));
}

void HeatSolver::vcycle_host(const std::vector<double> &r, std::vector<double> &z) {
    Impl &I = *d;
    const long long n = I.ist(0) * I.nsys;
    if ((long long)r.size() != n) throw std::runtime_error("r has the wrong size");
    HCHECK(hipMemcpy(I.q, r.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    double *zp = I.Z[0];
    I.vcycle(I.q, zp);
    std::vector<double> pad(I.pst(0) * I.nsys);
    HCHECK(hipStreamSynchronize(I.stream));
    HCHECK(hipMemcpy(pad.data(), zp, sizeof(double) * pad.size(), hipMemcpyDeviceToHost));
    const int m = I.L[0].m;
    z.resize(n);
    for (int s = 0; s < I.nsys; s++)
        for (int bb = 0; bb < I.nb; bb++)
            for (int c = 0; c < m * m * m; c++) {
                const int i = c / (m * m), j = (c / m) % m, k = c % m;
                z[s * I.ist(0) + (long long)bb * m * m * m + c] =
                    pad[s * I.pst(0) + (long long)bb * I.L[0].P + ((long long)(i + 1) * (m + 2) + (j + 1)) * (m + 2) + k + 1];
            }
}

HeatResult HeatSolver::solve(double tol, int max_iter, bool use_corrector) {
    return solve(tol, max_iter, use_corrector ? kSolveLearned : kSolvePlain, nullptr);
}

void HeatSolver::Impl::ensure_learned() {
    if (Az) return;
    Az = dalloc(ist(0) * nsys);
    rp = dalloc(ist(0) * nsys);
    Ac = dalloc(ist(0) * nsys);
    cpad = dalloc(pst(0) * nsys);
    partial_cells = dalloc((size_t)nsys * 3 * ncell_blocks());
    void *a = nullptr;
    HCHECK(hipMalloc(&a, sizeof(float) * ist(0) * nsys));
    allocs.push_back(a);
    xin = (float *)a;
}

// Az = A z, r' = r - A z, network input r'/diag and its per-system scale
// (root mean square) in sc[3]
void HeatSolver::Impl::features(double *z) {
    exchange(0, z);
    hipLaunchKernelGGL(k_hfeat, grid_cells(0), dim3(256), 0, stream, L[0], (const double *)z, (const double *)r, Az, rp, xin,
                       (const double *)inv_dt, partial_cells, pst(0), ist(0));
    final_dots(1, ncell_blocks());
    hipLaunchKernelGGL(k_hscalars, dim3((nsys + 63) / 64), dim3(64), 0, stream, 3, nsys, (const double *)dots, sc,
                       (const int *)active, (double)ist(0));
}

// c = scale * net(features / scale), a = <c,r'>/<c,Ac>, z += a c, Az += a Ac,
// and beta = <z, q>/gamma of the FCG direction update (sc[1])
void HeatSolver::Impl::learned_correction(double *z) {
    corr->infer_heat(xin, sc + 3, kSc, nb, pst(0), cpad, (long long)nsys * nb, stream);
    exchange(0, cpad);
    hipLaunchKernelGGL(k_happly_dot, grid_cells(0), dim3(256), 0, stream, L[0], (const double *)cpad, (const double *)rp,
                       Ac, (const double *)inv_dt, partial_cells, pst(0), ist(0));
    final_dots(2, ncell_blocks());
    hipLaunchKernelGGL(k_hscalars, dim3((nsys + 63) / 64), dim3(64), 0, stream, 2, nsys, (const double *)dots, sc,
                       (const int *)active, 0.0);
    hipLaunchKernelGGL(k_hzcorr, grid_cells(0), dim3(256), 0, stream, L[0], z, (const double *)cpad, Az,
                       (const double *)Ac, (const double *)q, (const double *)sc, partial_cells, pst(0), ist(0));
    final_dots(1, ncell_blocks());
    hipLaunchKernelGGL(k_hscalars, dim3((nsys + 63) / 64), dim3(64), 0, stream, 1, nsys, (const double *)dots, sc,
                       (const int *)active, 0.0);
}

HeatResult HeatSolver::solve(double tol, int max_iter, int mode, HeatRecord *rec) {
    Impl &I = *d;
    const int ns = I.nsys;
    const long long ni = I.ist(0);
    const bool learned = mode == kSolveLearned;
    if (learned && !I.corr) throw std::runtime_error("no learned correction is attached to the heat solver");
    if ((learned || mode == kSolveRecord) && I.L[0].m != 16) throw std::runtime_error("the learned correction needs 16^3 blocks");
    if (mode == kSolveRecord && !rec) throw std::runtime_error("record mode needs a record buffer");
    if (learned || mode == kSolveRecord) I.ensure_learned();
    HeatResult res;
    res.iters.assign(ns, 0);
    res.relres.assign(ns, 0.0);
    std::vector<int> act(ns, 1);
    HCHECK(hipMemcpy(I.active, act.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HCHECK(hipStreamSynchronize(I.stream));
    const auto t0 = std::chrono::steady_clock::now();
    const dim3 g0 = I.grid_cells(0);
    const long long rec0 = rec ? rec->count : 0;
    // x = 0, r = b
    I.zero(I.x, ni * ns);
    hipLaunchKernelGGL(k_hcopy, dim3((unsigned)((ni * ns + 255) / 256)), dim3(256), 0, I.stream, (const double *)I.b, I.r,
                       ni * ns);
    I.dot(0, 1, I.b, 0, I.b, 0);
    hipLaunchKernelGGL(k_hscalars, dim3((ns + 63) / 64), dim3(64), 0, I.stream, 4, ns, (const double *)I.dots, I.sc,
                       (const int *)I.active, 0.0);
    std::vector<double> bn(3 * ns), rn(3 * ns);
    HCHECK(hipMemcpyAsync(bn.data(), I.dots, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, I.stream));
    HCHECK(hipStreamSynchronize(I.stream));
    // a system with b = 0 has the solution 0: it never becomes active
    std::vector<int> done(ns, 0);
    for (int s = 0; s < ns; s++)
        if (bn[3 * s] == 0.0) { done[s] = 1; act[s] = 0; }
    HCHECK(hipMemcpy(I.active, act.data(), sizeof(int) * ns, hipMemcpyHostToDevice));

    auto record = [&](const double *z) {
        // one snapshot (residual/error pair) per system still iterating, each
        // contributing per_snapshot random blocks, while room is left
        std::vector<int> rs, rb;
        for (int s = 0; s < ns; s++) {
            if (done[s]) continue;
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
                           rec->W + rec->count * 4096, (const double *)I.sc, rec->scale + rec->count, I.pst(0), ni);
        HCHECK(hipStreamSynchronize(I.stream));   // rs/rb host vectors leave scope
        rec->count += R;
    };
    auto precond = [&]() -> double * {
        double *z = I.Z[0];
        if (learned) {
            I.vcycle(I.r, z, I.learned_pre0, I.learned_post0);
            I.features(z);
            I.learned_correction(z);        // also beta for the direction update
        } else if (mode == kSolveRecord) {
            I.vcycle(I.r, z, I.learned_pre0, I.learned_post0);
            I.features(z);
            record(z);
        } else {
            I.vcycle(I.r, z);
        }
        return z;
    };
    double *z = precond();
    if (learned)
        hipLaunchKernelGGL(k_hupdate_pq, g0, dim3(256), 0, I.stream, I.L[0], I.p, (const double *)z, I.q,
                           (const double *)I.Az, (const double *)I.sc, I.pst(0), ni, 0);
    else
        hipLaunchKernelGGL(k_hupdate_p, g0, dim3(256), 0, I.stream, I.L[0], I.p, (const double *)z, (const double *)I.sc,
                           I.pst(0), 0);
    int it = 0;
    bool all0 = true;
    for (int s = 0; s < ns; s++) all0 = all0 && done[s];

    if (mode != kSolveRecord && I.use_graph) {
        // One FCG iteration captured as a hipGraph and replayed; the convergence
        // test runs on the device (k_hconverge), the host reads one counter per
        // iteration.  Arithmetic and iteration counts equal the eager loop's.
        auto body = [&]() {
            if (!learned) {
                I.exchange(0, I.p);
                hipLaunchKernelGGL((k_happly<0>), g0, dim3(256), 0, I.stream, I.L[0], (const double *)I.p,
                                   (const double *)nullptr, I.q, (const double *)I.inv_dt, I.pst(0), ni, 0.0);
            }
            I.dot(0, 2, I.p, 1, I.q, 0, I.p, 1, I.r, 0);
            hipLaunchKernelGGL(k_hscalars, dim3((ns + 63) / 64), dim3(64), 0, I.stream, 0, ns, (const double *)I.dots,
                               I.sc, (const int *)I.active, 0.0);
            hipLaunchKernelGGL(k_hupdate_xr, g0, dim3(256), 0, I.stream, I.L[0], I.x, I.r, (const double *)I.p,
                               (const double *)I.q, (const double *)I.sc, I.pst(0), ni);
            I.dot(0, 1, I.r, 0, I.r, 0);
            hipLaunchKernelGGL(k_hconverge, dim3(1), dim3(256), 0, I.stream, ns, (const double *)I.dots,
                               (const double *)I.sc, tol, I.active, I.iters_dev, I.relres_dev, I.nactive_dev);
            double *zz = precond();
            if (learned) {
                hipLaunchKernelGGL(k_hupdate_pq, g0, dim3(256), 0, I.stream, I.L[0], I.p, (const double *)zz, I.q,
                                   (const double *)I.Az, (const double *)I.sc, I.pst(0), ni, 1);
            } else {
                I.dot(0, 1, zz, 1, I.q, 0);
                hipLaunchKernelGGL(k_hscalars, dim3((ns + 63) / 64), dim3(64), 0, I.stream, 1, ns,
                                   (const double *)I.dots, I.sc, (const int *)I.active, 0.0);
                hipLaunchKernelGGL(k_hupdate_p, g0, dim3(256), 0, I.stream, I.L[0], I.p, (const double *)zz,
                                   (const double *)I.sc, I.pst(0), 1);
            }
        };
        hipLaunchKernelGGL(k_hizero, dim3((ns + 255) / 256), dim3(256), 0, I.stream, I.iters_dev, ns);
        if (learned) I.corr->prepare(I.stream);
        hipGraphExec_t &G = I.graph[learned ? 1 : 0];
        if (!G) {
            hipGraph_t g;
            HCHECK(hipStreamBeginCapture(I.stream, hipStreamCaptureModeThreadLocal));
            body();
            HCHECK(hipStreamEndCapture(I.stream, &g));
            HCHECK(hipGraphInstantiate(&G, g, nullptr, nullptr, 0));
            HCHECK(hipGraphDestroy(g));
        }
        int nact = 0;
        for (int s = 0; s < ns; s++) nact += done[s] ? 0 : 1;
        int launched = 0;
        while (nact > 0 && launched < max_iter) {
            HCHECK(hipGraphLaunch(G, I.stream));
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
            res.iters[s] = done[s] ? 0 : itd[s];
            res.relres[s] = done[s] ? 0.0 : rel[s];
        }
        HCHECK(hipGetLastError());
        return res;
    }

    while (!all0) {
        if (!learned) {
            I.exchange(0, I.p);
            hipLaunchKernelGGL((k_happly<0>), g0, dim3(256), 0, I.stream, I.L[0], (const double *)I.p,
                               (const double *)nullptr, I.q, (const double *)I.inv_dt, I.pst(0), ni, 0.0);
        }
        I.dot(0, 2, I.p, 1, I.q, 0, I.p, 1, I.r, 0);
        hipLaunchKernelGGL(k_hscalars, dim3((ns + 63) / 64), dim3(64), 0, I.stream, 0, ns, (const double *)I.dots, I.sc,
                           (const int *)I.active, 0.0);
        hipLaunchKernelGGL(k_hupdate_xr, g0, dim3(256), 0, I.stream, I.L[0], I.x, I.r, (const double *)I.p,
                           (const double *)I.q, (const double *)I.sc, I.pst(0), ni);
        it++;
        I.dot(0, 1, I.r, 0, I.r, 0);
        HCHECK(hipMemcpyAsync(rn.data(), I.dots, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, I.stream));
        HCHECK(hipStreamSynchronize(I.stream));
        bool all = true, changed = false;
        for (int s = 0; s < ns; s++) {
            if (done[s]) continue;
            const double rel = std::sqrt(rn[3 * s] / bn[3 * s]);
            res.relres[s] = rel;
            res.iters[s] = it;
            if (rel <= tol || !std::isfinite(rel)) { done[s] = 1; act[s] = 0; changed = true; }
            else all = false;
        }
        if (changed) HCHECK(hipMemcpyAsync(I.active, act.data(), sizeof(int) * ns, hipMemcpyHostToDevice, I.stream));
        if (all || it >= max_iter) break;
        z = precond();
        if (learned) {
            hipLaunchKernelGGL(k_hupdate_pq, g0, dim3(256), 0, I.stream, I.L[0], I.p, (const double *)z, I.q,
                               (const double *)I.Az, (const double *)I.sc, I.pst(0), ni, 1);
        } else {
            I.dot(0, 1, z, 1, I.q, 0);
            hipLaunchKernelGGL(k_hscalars, dim3((ns + 63) / 64), dim3(64), 0, I.stream, 1, ns, (const double *)I.dots,
                               I.sc, (const int *)I.active, 0.0);
            hipLaunchKernelGGL(k_hupdate_p, g0, dim3(256), 0, I.stream, I.L[0], I.p, (const double *)z,
                               (const double *)I.sc, I.pst(0), 1);
        }
    }
    HCHECK(hipStreamSynchronize(I.stream));
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
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

void HeatSolver::vcycle_residual(const double *r_dev, double *z_dev_padded, double *resid_dev) {
    Impl &I = *d;
    double *z = I.Z[0];
    I.vcycle(r_dev, z);
    I.exchange(0, z);
    hipLaunchKernelGGL((k_happly<1>), I.grid_cells(0), dim3(256), 0, I.stream, I.L[0], (const double *)z, r_dev, resid_dev,
                       (const double *)I.inv_dt, I.pst(0), I.ist(0), 0.0);
    if (z_dev_padded)
        hipLaunchKernelGGL(k_hcopy, dim3((unsigned)((I.pst(0) * I.nsys + 255) / 256)), dim3(256), 0, I.stream,
                           (const double *)z, z_dev_padded, I.pst(0) * I.nsys);
}

double *HeatSolver::device_solution() { return d->x; }
double *HeatSolver::device_rhs() { return d->b; }
const HLevView HeatSolver::level0_view() const {
    HLevView v;
    v.m = d->L[0].m; v.nb = d->nb; v.P = d->L[0].P; v.M3 = d->L[0].M3; v.nsys = d->nsys;
    return v;
}

}  // namespace falcon
