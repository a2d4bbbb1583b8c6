// Handle and C-ABI of libfalcon_hip.so (declarations in include/falcon_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/falcon_hip.h"
#include "device.hpp"
#include "cnn.hpp"
#include "heat.hpp"
#include "layout.hpp"
#include "output.hpp"

// Scratch of falcon_compute_dt_noalloc, in device memory the library never
// allocates: the dt partial minima (one per tile of the diffusion's edge, at
// most DT_NOALLOC_MAX_PARTIALS) followed by the resulting dt.  Each GPU holds its
// own copy until the process ends; every call on one GPU shares it.
constexpr int DT_NOALLOC_MAX_PARTIALS = 1 << 14;
static __device__ __align__(256) double scratch[DT_NOALLOC_MAX_PARTIALS + 32];

namespace falcon {
void synthetic_modes(const Tree &t, uint64_t seed, std::vector<double> &modes);
}

using namespace falcon;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHECK(x)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) throw Error(std::string(#x) + ": " + hipGetErrorString(e_));            \
    } while (0)

template <typename F>
int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// Device address of `scratch` per GPU, looked up once per device at set-up
// (falcon_set_gc_all) with that device current; host code uses these, never the
// array's name.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

void scratch_setup(int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) throw Error("device id out of range for the dt scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return;
    void *p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error("hipGetSymbolAddress returned no address for the dt scratch");
    g_scratch_ptr[dev] = (double *)p;
}

int host_threads() {
    unsigned n = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(n, 32u));
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void alloc(size_t n) {
        free();
        if (n) HIPCHECK(hipMalloc(&p, n));
        bytes = n;
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const { return (T *)p; }
};

}  // namespace

struct falcon_handle {
    Layout lay;
    bool labelled = false, built = false, tables = false;
    bool bg_dirty = false;    // a background changed after the device tables went up
    std::vector<double> bg;   // per block 8 x NI
    std::vector<char> bg_set;
    int precision = 0;
    int device = -1;           // the GPU current at falcon_set_gc_all (the stream's device)
    hipStream_t stream = nullptr;
    Dims dims{};
    DevTables dt{};
    DevGhosts gh64{}, gh32{};   // FP64 buffers (exact) and the FP32 deviation path
    DevBoundary bnd{};
    // device memory
    DevBuf meta, radial, angular, ghost64[4], ghost32[4], ghost_copy32, ghost_lerp32[2], bnd_list;   // ghost tables
    DevBuf ghost_rest32, ghost_pull32;   // copies not read in place, pull table (FP32 path)
    bool ghost_pull = true;              // FP32 step: kernels read same-level ghost copies in place
    bool device_tables = false;          // falcon_set_gc_all builds the ghost tables on the device

    DevBuf U[2], F32[2], S64[2], S32[2], dtp, partial, rbuf;
    // GLM divergence cleaning (falcon_set_divb, FP64 path): psi beside U[] and
    // the stage buffers, one padded variable per block, and grad(psi) of the
    // current stage
    int divb = 0;
    DevBuf psi[2], psiS[2], glmG, glmC;
    DevBuf diffbuf;          // falcon_state_diff's reduction words
    DevBuf scanbuf;          // falcon_check_state's count and first position
    int cur = 0;             // index of the current state buffer in U[]
    // validity of derived device data
    bool partials_valid = false;   // partial[] holds min(h/v_wave) of U[cur]
    bool shadow_valid = false;     // F32[cur] (with ghosts) is the FP32 copy of U[cur]
    bool u_ghosts_valid = true;    // U[cur]'s ghost cells are filled
    // profiling: (kind, start event, end event) per launch when enabled
    std::vector<hipEvent_t> *prof_ev = nullptr;
    std::vector<int> *prof_kind = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_cur = -1, graph_prec = -1, graph_divb = -1;
    // launch shape of the stage and diffusion marches per precision (0 FP64,
    // 1 FP32): (theta, phi) tile edge, 8 or 16, and workgroups (0: one per
    // tile; n > 0: n workgroups looping over the tiles); falcon_autotune picks
    // them by timing
    int tile_stage[2] = {16, 16}, tile_diff[2] = {16, 16};
    int wgs_stage[2] = {0, 0}, wgs_diff[2] = {0, 0};
    std::unique_ptr<HeatSolver> heat;
    HeatParams heat_params;
    // PFSS form of the implicit solver (falcon_pfss_setup): its tables and
    // the magnetogram of the right-hand side (per bottom block, j fastest)
    bool pfss_on = false;
    PfssTables pfss;
    std::vector<double> pfss_br;
    std::unique_ptr<CnnModel> cnn;
    double graph_cfl = -1;
    // dt > 0: every step uses it in place of the CFL dt (falcon_set_fixed_dt)
    double fixed_dt = 0, graph_fixed_dt = 0;

    ~falcon_handle() {
        if (graph) (void)hipGraphExecDestroy(graph);
        DevBuf *all[] = {&meta, &radial, &angular, &ghost64[0], &ghost64[1], &ghost64[2], &ghost64[3], &ghost32[0],
                         &ghost32[1], &ghost32[2], &ghost32[3], &ghost_copy32, &ghost_lerp32[0], &ghost_lerp32[1],
                         &bnd_list, &ghost_rest32, &ghost_pull32,

                         &U[0], &U[1], &F32[0], &F32[1], &S64[0], &S64[1], &S32[0], &S32[1], &dtp, &partial, &rbuf, &diffbuf, &scanbuf,
                         &psi[0], &psi[1], &psiS[0], &psiS[1], &glmG, &glmC};
        for (DevBuf *b : all) b->free();
        if (stream) (void)hipStreamDestroy(stream);
    }

    int NI() const { return lay.tree.ni + 2 * lay.tree.ng; }
    int NJ() const { return lay.tree.nj + 2 * lay.tree.ng; }
    int NK() const { return lay.tree.nk + 2 * lay.tree.ng; }

    void need_label() {
        if (!labelled) {
            lay.tree.label();
            labelled = true;
            bg.assign((size_t)lay.tree.nleaf * 8 * NI(), 0.0);
            bg_set.assign(lay.tree.nleaf, 0);
        }
    }
    void need_tables() const {
        if (!tables) throw Error("falcon_set_gc_all has not been called");
    }
    void need_built() const {
        if (!built) throw Error("falcon_build_tables has not been called");
    }
    void check_block(int ib) const {
        if (ib < 0 || ib >= lay.tree.nleaf) throw Error("block index out of range");
    }
};

namespace {

// host Fortran (i fastest) <-> device block variables (cell_off) for one block
void fortran_to_device(const double *src, double *dst, const Dims &d) {
    const size_t P = (size_t)d.P;
    for (int v = 0; v < 8; v++)
        for (int K = 0; K < d.NK; K++)
            for (int J = 0; J < d.NJ; J++)
                for (int I = 0; I < d.NI; I++)
                    dst[v * P + cell_off(d, I, J, K)] = src[v * P + (size_t)I + (size_t)d.NI * (J + (size_t)d.NJ * K)];
}
void device_to_fortran(const double *src, double *dst, const Dims &d) {
    const size_t P = (size_t)d.P;
    for (int v = 0; v < 8; v++)
        for (int K = 0; K < d.NK; K++)
            for (int J = 0; J < d.NJ; J++)
                for (int I = 0; I < d.NI; I++)
                    dst[v * P + (size_t)I + (size_t)d.NI * (J + (size_t)d.NJ * K)] = src[v * P + cell_off(d, I, J, K)];
}

// FP32 buffers of the deviation path: same-panel ghost entries whose weights
// lie within 1e-12 of 0 or 1 (same-level neighbours whose cell centres
// coincide up to rounding) are plain copies from the nearest donor cell;
// same-panel entries with only the r weight away from 0/1 interpolate along r.
bool fp32_near(const GhostEntry &g) {
    if (g.yy) return false;
    for (int a = 0; a < 3; a++)
        if (!(g.w[a] < 1e-12 || g.w[a] > 1.0 - 1e-12)) return false;
    return true;
}
bool fp32_lerp_r(const GhostEntry &g) {
    auto snapped = [](double w) { return w < 1e-12 || w > 1.0 - 1e-12; };
    return !g.yy && !snapped(g.w[0]) && snapped(g.w[1]) && snapped(g.w[2]);
}

// (dst, src) of the copies and (dst, base) with the r weight of the radial
// lerps, each sorted by destination; offsets block*8*P + padded cell.
void fp32_copy_lerp(const falcon_handle *h, std::vector<int2> &copy32, std::vector<std::pair<int2, double>> &lerp32) {
    const Dims &d = h->dims;
    const int ng = d.ng;
    // Fortran (1-based interior) index -> cell offset
    auto cell = [&](int i, int j, int k) -> long long { return cell_off(d, i + ng - 1, j + ng - 1, k + ng - 1); };
    copy32.clear();
    lerp32.clear();
    for (const GhostEntry &g : h->lay.ghosts) {
        const int32_t d0 = (int32_t)((long long)g.recv_block * 8 * d.P + cell(g.gi, g.gj, g.gk));
        // nearest donor cell along theta and phi; along r too for a copy
        const int jn = g.j0 + (g.w[1] > 0.5 ? 1 : 0), kn = g.k0 + (g.w[2] > 0.5 ? 1 : 0);
        const long long s0 = (long long)g.donor * 8 * d.P + cell(g.i0, jn, kn);
        if (fp32_near(g))
            copy32.push_back(make_int2(d0, (int32_t)((long long)g.donor * 8 * d.P +
                                                     cell(g.i0 + (g.w[0] > 0.5 ? 1 : 0), jn, kn))));
        else if (fp32_lerp_r(g)) lerp32.push_back({make_int2(d0, (int32_t)s0), g.w[0]});
    }
    std::sort(copy32.begin(), copy32.end(), [](const int2 &a, const int2 &b) { return a.x < b.x; });
    std::sort(lerp32.begin(), lerp32.end(), [](const auto &a, const auto &b) { return a.first.x < b.first.x; });
}

// Pull table of the FP32 step.  A copy entry whose ghost (padded index g) and
// donor cell (padded index p, interior) differ by 0 or +-n per axis is a
// same-level translation; seen from the donor D the receiver lies in direction
// dir = -(g - p)/n.  (D, dir) qualifies when all ghosts of the image of D's
// band for dir (ng wide along every axis with dir != 0, n along the others)
// are such entries with one receiving block R: then R reads its ghosts in
// direction -dir from D, pull[R*27 + code(-dir)] = D (code(d) = (d0+1)*9 +
// (d1+1)*3 + (d2+1); -1 elsewhere), and those entries leave the copies the
// step runs (rest, in copy order).
void build_pull_table(const falcon_handle *h, const std::vector<int2> &copy32, std::vector<int32_t> &pull,
                      std::vector<int2> &rest) {
    const Tree &t = h->lay.tree;
    const Dims &d = h->dims;
    const int nb = t.nleaf, ng = t.ng;
    const long long P = d.P;
    const int n3[3] = {t.ni, t.nj, t.nk};
    const bool fits = n3[0] >= 2 * ng && n3[1] >= 2 * ng && n3[2] >= 2 * ng;
    std::vector<int32_t> recv((size_t)nb * 27, -1), cnt((size_t)nb * 27, 0);
    std::vector<char> bad((size_t)nb * 27, 0);
    auto decode = [&](long long off, int &blk, int p[3]) {
        blk = (int)(off / (8 * P));
        cell_ijk(d, off - (long long)blk * 8 * P, p[0], p[1], p[2]);
    };
    // donor-side direction key of a translation entry, or -1
    auto key_of = [&](const int2 &e, int &R) -> long long {
        int D, pg[3], pd[3], dir[3];
        decode(e.x, R, pg);
        decode(e.y, D, pd);
        for (int a = 0; a < 3; a++) {
            if (pd[a] < ng || pd[a] >= ng + n3[a]) return -1;
            const int o = pg[a] - pd[a];
            if (o == 0) dir[a] = 0;
            else if (o == n3[a]) dir[a] = -1;
            else if (o == -n3[a]) dir[a] = 1;
            else return -1;
        }
        if (!dir[0] && !dir[1] && !dir[2]) return -1;
        return (long long)D * 27 + (dir[0] + 1) * 9 + (dir[1] + 1) * 3 + (dir[2] + 1);
    };
    if (fits) {
        for (const int2 &e : copy32) {
            int R;
            const long long k = key_of(e, R);
            if (k < 0) continue;
            if (recv[k] == -1) recv[k] = R;
            else if (recv[k] != R) bad[k] = 1;
            cnt[k]++;
        }
    }
    std::vector<int32_t> by_donor((size_t)nb * 27, -1);
    pull.assign((size_t)nb * 27, -1);
    for (long long k = 0; k < (long long)nb * 27; k++) {
        const int code = (int)(k % 27);
        const int dir[3] = {code / 9 - 1, (code / 3) % 3 - 1, code % 3 - 1};
        long long band = 1;
        for (int a = 0; a < 3; a++) band *= dir[a] ? ng : n3[a];
        if (fits && recv[k] >= 0 && !bad[k] && cnt[k] == band) {
            by_donor[k] = recv[k];
            pull[(size_t)recv[k] * 27 + (26 - code)] = (int32_t)(k / 27);
        }
    }
    rest.clear();
    for (const int2 &e : copy32) {
        int R;
        const long long k = fits ? key_of(e, R) : -1;
        if (k < 0 || by_donor[k] != R) rest.push_back(e);
    }
}

constexpr int N_RADIAL_ROWS = 11, N_ANGULAR_ROWS = 3;

// Radial rows of every block from its 1-D background, as [row][block][I]:
// the 3_6/ModBlock.f90:255-262 products, then the per-term products of
// ModEquation.
std::vector<double> radial_rows(const falcon_handle *h) {
    const Tree &t = h->lay.tree;
    const int nb = t.nleaf, NI = h->NI(), NR = N_RADIAL_ROWS;
    std::vector<double> radT((size_t)nb * NR * NI);
    for (int b = 0; b < nb; b++) {
        if (!h->bg_set[b]) throw Error("background not set for block " + std::to_string(b));
        const BlockGeom &g = h->lay.geom[b];
        const double *B = h->bg.data() + (size_t)b * 8 * NI;
        for (int I = 0; I < NI; I++) {
            const double gg = B[I], rho0 = B[I + NI], p0 = B[I + 2 * NI], T0 = B[I + 3 * NI];
            const double g1 = B[I + 4 * NI], g3 = B[I + 5 * NI], heat = B[I + 6 * NI], xr = B[I + 7 * NI];
            const double p0r = p0 / rho0, rho0T0 = rho0 * T0, gm1 = g3 - 1.0, x = g.xi[I];
            auto row = [&](int a) -> double & { return radT[((size_t)a * nb + b) * NI + I]; };
            row(0) = x;
            row(1) = 1.0 / x;
            row(2) = 1.0 / (x * x);
            row(3) = rho0;
            row(4) = g1 * p0r;
            row(5) = gm1 * rho0T0;
            row(6) = gg / rho0;
            row(7) = -1.0 / (xr * xr);
            row(8) = (1.0 * (heat + 0.0)) / rho0T0;
            row(9) = (1.0 / xr) * std::sqrt(g1 * p0r);
            row(10) = std::sqrt(rho0);
        }
    }
    return radT;
}

// A background set after the tables went up: the radial rows are rewritten in
// place (same buffer, so captured graphs stay valid) and the dt partial minima,
// which depend on the sound speed, are recomputed before the next step.
void upload_radial(falcon_handle *h) {
    const std::vector<double> radT = radial_rows(h);
    if (h->radial.bytes != sizeof(double) * radT.size()) throw Error("radial table size changed");
    if (h->stream) HIPCHECK(hipStreamSynchronize(h->stream));
    HIPCHECK(hipMemcpy(h->radial.p, radT.data(), sizeof(double) * radT.size(), hipMemcpyHostToDevice));
    h->partials_valid = false;
    h->shadow_valid = false;
}

void upload_tables(falcon_handle *h) {
    const Tree &t = h->lay.tree;
    const int nb = t.nleaf, NI = h->NI(), NJ = h->NJ(), NK = h->NK(), ng = t.ng;
    // per-block constants
    std::vector<BlockMeta> meta(nb);
    const int NA = N_ANGULAR_ROWS;
    std::vector<double> radT = radial_rows(h), ang((size_t)nb * NA * NJ);
    for (int b = 0; b < nb; b++) {
        const BlockGeom &g = h->lay.geom[b];
        const double c12 = (double)(1.0f / 12.0f);
        meta[b] = {c12 / g.dxi, c12 / g.dxj, c12 / g.dxk, g.dxi, g.dxj, g.dxk, (g.bottom ? 1 : 0) | (g.top ? 2 : 0)};
        double *arow = ang.data() + (size_t)b * NA * NJ;
        for (int J = 0; J < NJ; J++) {
            double s, cot;
            sin_and_cot(g.xj[J], s, cot);
            arow[J] = s;
            arow[NJ + J] = 1.0 / s;
            arow[2 * NJ + J] = cot;
        }
    }
    // angular rows as [array][block][index] for simple pointers
    std::vector<double> angT(ang.size());
    for (int b = 0; b < nb; b++)
        for (int a = 0; a < NA; a++)
            std::memcpy(&angT[((size_t)a * nb + b) * NJ], &ang[((size_t)b * NA + a) * NJ], sizeof(double) * NJ);
    h->meta.alloc(sizeof(BlockMeta) * nb);
    h->radial.alloc(sizeof(double) * radT.size());
    h->angular.alloc(sizeof(double) * angT.size());
    HIPCHECK(hipMemcpy(h->meta.p, meta.data(), sizeof(BlockMeta) * nb, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h->radial.p, radT.data(), sizeof(double) * radT.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h->angular.p, angT.data(), sizeof(double) * angT.size(), hipMemcpyHostToDevice));
    const double *R = h->radial.as<double>(), *A = h->angular.as<double>();
    const size_t rs = (size_t)nb * NI, as = (size_t)nb * NJ;
    DevTables &dt = h->dt;
    dt.meta = h->meta.as<BlockMeta>();
    dt.xi = R; dt.rinv = R + rs; dt.r2inv = R + 2 * rs; dt.rho0 = R + 3 * rs; dt.gp = R + 4 * rs; dt.gq = R + 5 * rs;
    dt.g_over_rho0 = R + 6 * rs; dt.xir2 = R + 7 * rs; dt.heat = R + 8 * rs; dt.cs = R + 9 * rs; dt.sqrt_rho0 = R + 10 * rs;
    dt.sint = A; dt.sinv = A + as; dt.cot = A + 2 * as;

    // ghost entries: same-panel first, then Yin-Yang, each in target-list order
    const auto &G = h->lay.ghosts;
    const long long P = h->dims.P;
    auto cell = [&](int i, int j, int k) -> long long { return cell_off(h->dims, i + ng - 1, j + ng - 1, k + ng - 1); };
    if ((long long)nb * 8 * P >= (1ll << 31)) throw Error("state too large for 32-bit ghost offsets");
    // FP64 buffers: every entry weighted, same-panel entries first, then
    // Yin-Yang, each in target-list order.  FP32 buffers of the deviation path:
    // near copies and radial lerps (fp32_copy_lerp), the rest weighted in the
    // same order.
    std::vector<int2> copy32;
    std::vector<std::pair<int2, double>> lerp32;
    fp32_copy_lerp(h, copy32, lerp32);
    std::vector<int2> lerp_idx(lerp32.size());
    std::vector<double> lerp_w(lerp32.size());
    for (size_t q = 0; q < lerp32.size(); q++) { lerp_idx[q] = lerp32[q].first; lerp_w[q] = lerp32[q].second; }
    auto up = [&](DevBuf &buf, const void *data, size_t bytes) {
        buf.alloc(std::max<size_t>(bytes, 16));
        if (bytes) HIPCHECK(hipMemcpy(buf.p, data, bytes, hipMemcpyHostToDevice));
    };
    // weighted table of the entries passing `keep`, plain then Yin-Yang
    auto weighted = [&](auto keep, DevBuf *bufs) -> DevGhosts {
        std::vector<size_t> order;
        for (int pass = 0; pass < 2; pass++)
            for (size_t e = 0; e < G.size(); e++)
                if (G[e].yy == pass && keep(G[e])) order.push_back(e);
        const size_t NG = order.size();
        int yy_begin = 0;
        while (yy_begin < (int)NG && !G[order[yy_begin]].yy) yy_begin++;
        std::vector<int32_t> dst(NG), src(NG);
        std::vector<double> w(3 * NG), rot(2 * NG);
        for (size_t q = 0; q < NG; q++) {
            const GhostEntry &e = G[order[q]];
            dst[q] = (int32_t)((long long)e.recv_block * 8 * P + cell(e.gi, e.gj, e.gk));
            src[q] = (int32_t)((long long)e.donor * 8 * P + cell(e.i0, e.j0, e.k0));
            w[3 * q] = e.w[0]; w[3 * q + 1] = e.w[1]; w[3 * q + 2] = e.w[2];
            rot[2 * q] = e.rot[0]; rot[2 * q + 1] = e.rot[1];
        }
        up(bufs[0], dst.data(), sizeof(int32_t) * NG);
        up(bufs[1], src.data(), sizeof(int32_t) * NG);
        up(bufs[2], w.data(), sizeof(double) * 3 * NG);
        up(bufs[3], rot.data(), sizeof(double) * 2 * NG);
        return DevGhosts{bufs[0].as<int32_t>(), bufs[1].as<int32_t>(), bufs[2].as<double>(), bufs[3].as<double>(),
                         0, (int)NG, yy_begin, nullptr, 0, nullptr, nullptr, 0};
    };
    h->gh64 = weighted([](const GhostEntry &) { return true; }, h->ghost64);
    h->gh32 = weighted([&](const GhostEntry &g) { return !fp32_near(g) && !fp32_lerp_r(g); }, h->ghost32);
    up(h->ghost_copy32, copy32.data(), sizeof(int2) * copy32.size());
    up(h->ghost_lerp32[0], lerp_idx.data(), sizeof(int2) * lerp_idx.size());
    up(h->ghost_lerp32[1], lerp_w.data(), sizeof(double) * lerp_w.size());
    h->gh32.copy = h->ghost_copy32.as<int2>();
    h->gh32.n_copy = (int)copy32.size();
    h->gh32.lerp = h->ghost_lerp32[0].as<int2>();
    h->gh32.lerp_w = h->ghost_lerp32[1].as<double>();
    h->gh32.n_lerp = (int)lerp_idx.size();

    {
        std::vector<int32_t> pull;
        std::vector<int2> rest;
        build_pull_table(h, copy32, pull, rest);
        up(h->ghost_rest32, rest.data(), sizeof(int2) * rest.size());
        up(h->ghost_pull32, pull.data(), sizeof(int32_t) * pull.size());
        h->gh32.copy_rest = h->ghost_rest32.as<int2>();
        h->gh32.n_copy_rest = (int)rest.size();
        h->gh32.pull = h->ghost_pull32.as<int32_t>();
    }

    std::vector<int32_t> bl;
    for (int b = 0; b < nb; b++) {
        if (h->lay.geom[b].bottom) bl.push_back(b * 2 + 0);
        if (h->lay.geom[b].top) bl.push_back(b * 2 + 1);
    }
    h->bnd_list.alloc(sizeof(int32_t) * std::max<size_t>(bl.size(), 1));
    if (!bl.empty()) HIPCHECK(hipMemcpy(h->bnd_list.p, bl.data(), sizeof(int32_t) * bl.size(), hipMemcpyHostToDevice));
    h->bnd = {h->bnd_list.as<int32_t>(), (int)bl.size()};
}

void alloc_state(falcon_handle *h) {
    const size_t n = (size_t)h->dims.nblocks * 8 * h->dims.P;
    for (int q = 0; q < 2; q++) {
        if (!h->U[q].p) {
            h->U[q].alloc(n * sizeof(double));
            HIPCHECK(hipMemset(h->U[q].p, 0, n * sizeof(double)));
        }
    }
    if (!h->dtp.p) { h->dtp.alloc(sizeof(double)); HIPCHECK(hipMemset(h->dtp.p, 0, sizeof(double))); }
    // one partial minimum per workgroup of the diffusion / k_dt_tile; the 8-edge
    // tiling has the most
    const long long nt8 = (long long)h->dims.nblocks * ((h->dims.nj + 7) / 8) * ((h->dims.nk + 7) / 8);
    if (!h->partial.p) h->partial.alloc(sizeof(double) * std::max(1LL, nt8));
}

void alloc_stage(falcon_handle *h) {
    const size_t n = (size_t)h->dims.nblocks * 8 * h->dims.P;
    for (int q = 0; q < 2; q++) {
        if (h->precision == 0 && !h->S64[q].p) {
            h->S64[q].alloc(n * sizeof(double));
            HIPCHECK(hipMemset(h->S64[q].p, 0, n * sizeof(double)));
        }
        if (h->precision == 1 && !h->S32[q].p) {
            h->S32[q].alloc(n * sizeof(float));
            HIPCHECK(hipMemset(h->S32[q].p, 0, n * sizeof(float)));
        }
        if (h->precision == 1 && !h->F32[q].p) {
            h->F32[q].alloc(n * sizeof(float));
            HIPCHECK(hipMemset(h->F32[q].p, 0, n * sizeof(float)));
        }
    }
}

enum Kind { K_DT, K_BC, K_STAGE, K_XSTAGE, K_XSTATE, K_DIFF, K_OTHER, K_NKIND };

template <typename F>
void timed(falcon_handle *h, int kind, F &&f) {
    if (!h->prof_ev) { f(); return; }
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    HIPCHECK(hipEventRecord(a, h->stream));
    f();
    HIPCHECK(hipEventRecord(b, h->stream));
    h->prof_ev->push_back(a);
    h->prof_ev->push_back(b);
    h->prof_kind->push_back(kind);
}

// One RK4 step: ModAdvance_rk4 (3_6/ModAdvance.f90:18-112).  Stage buffers are
// ping-ponged; the diffusion writes the other state buffer, which becomes U^{n+1}.
// The diffusion reads the state's radial ghosts as stage 1's boundary call left
// them (the reference applies boundaries only to each stage's input), and those
// rows are then copied into U^{n+1} so its ghosts hold what primitive_IV holds.
// dt comes from the partial minima the previous diffusion left (or k_dt_tile).
// FP32 mode: the stencil input of every stage and of the diffusion is FP32 (the
// stage buffers and an FP32 shadow of the state); U^n, the stage accumulation
// and U^{n+1} stay FP64; every ghost fill runs on FP32 data with FP64 weights.
void enqueue_step(falcon_handle *h, double cfl, int cur) {
    const Dims &d = h->dims;
    const DevTables &t = h->dt;
    hipStream_t s = h->stream;
    double *U = h->U[cur].as<double>(), *Un = h->U[1 - cur].as<double>();
    double *dtp = h->dtp.as<double>(), *part = h->partial.as<double>();
    const int ts = h->tile_stage[h->precision], td = h->tile_diff[h->precision];
    const int ws = h->wgs_stage[h->precision], wd = h->wgs_diff[h->precision];
    timed(h, K_DT, [&] {
        if (h->fixed_dt > 0) launch_set_scalar(dtp, h->fixed_dt, s);
        else launch_dt_final(d, part, dtp, cfl, s, td);
    });
    if (h->precision == 0) {
        double *S0 = h->S64[0].as<double>(), *S1 = h->S64[1].as<double>();
        const double *in[4] = {U, S0, S1, S0};
        double *out[4] = {S0, S1, S0, U};
        // GLM: psi follows the same buffer roles; its radial boundary ghosts
        // are never written (ModBoundary sets none), so every psi buffer
        // keeps the ones falcon_set_psi gave it
        const bool glm = h->divb == 1;
        double *P = glm ? h->psi[cur].as<double>() : nullptr, *Pn = glm ? h->psi[1 - cur].as<double>() : nullptr;
        double *PS0 = glm ? h->psiS[0].as<double>() : nullptr, *PS1 = glm ? h->psiS[1].as<double>() : nullptr;
        double *G = glm ? h->glmG.as<double>() : nullptr, *Ch = glm ? h->glmC.as<double>() : nullptr;
        const double *pin[4] = {P, PS0, PS1, PS0};
        double *pout[4] = {PS0, PS1, PS0, P};
        if (glm) timed(h, K_OTHER, [&] { launch_glm_speed(d, t, U, Ch, s); });
        for (int k = 1; k <= 4; k++) {
            const double div = (double)(5.0f - (float)k);
            timed(h, K_BC, [&] { launch_boundary<double>(d, h->bnd, const_cast<double *>(in[k - 1]), s); });
            if (glm)
                timed(h, K_OTHER,
                      [&] { launch_glm_stage(d, t, in[k - 1], Ch, pin[k - 1], P, pout[k - 1], G, dtp, div, s); });
            timed(h, K_STAGE, [&] {
                launch_stage<double, double, double, double>(d, t, in[k - 1], U, out[k - 1], nullptr, dtp, div, s,
                                                             nullptr, ts, ws, G);
            });
            timed(h, k < 4 ? K_XSTAGE : K_XSTATE, [&] {
                if (glm) launch_exchange_glm(d, h->gh64, out[k - 1], pout[k - 1], s);
                else launch_exchange<double>(d, h->gh64, out[k - 1], s);
            });
        }
        timed(h, K_DIFF, [&] { launch_diffusion<double, double>(d, t, U, U, Un, nullptr, part, dtp, s, nullptr, td, wd); });
        if (glm) timed(h, K_OTHER, [&] { launch_glm_diffusion(d, t, U, P, Pn, dtp, s); });
        timed(h, K_OTHER, [&] { launch_copy_radial<double>(d, h->bnd, U, Un, s); });
        timed(h, K_XSTATE, [&] {
            if (glm) launch_exchange_glm(d, h->gh64, Un, Pn, s);
            else launch_exchange<double>(d, h->gh64, Un, s);
        });
    } else {
        float *S0 = h->S32[0].as<float>(), *S1 = h->S32[1].as<float>();
        float *F = h->F32[cur].as<float>(), *Fn = h->F32[1 - cur].as<float>();
        const float *in[4] = {F, S0, S1, S0};
        float *out[3] = {S0, S1, S0};
        // with the pull table, the stage and diffusion kernels read same-level
        // ghost copies from the neighbours' interiors and the exchange fills
        // only the other ghosts
        const int32_t *pull = h->ghost_pull ? h->gh32.pull : nullptr;
        auto exchange = [&](float *st) {
            if (pull) launch_exchange_rest<float>(d, h->gh32, st, s);
            else launch_exchange<float>(d, h->gh32, st, s);
        };
        for (int k = 1; k <= 4; k++) {
            timed(h, K_BC, [&] { launch_boundary<float>(d, h->bnd, const_cast<float *>(in[k - 1]), s, pull); });
            const double div = (double)(5.0f - (float)k);
            if (k < 4) {
                // intermediate stage values are FP32 temporaries: base = FP32 copy of U^n
                timed(h, K_STAGE, [&] {
                    launch_stage<float, float, float, float>(d, t, in[k - 1], F, out[k - 1], nullptr, dtp, div, s, pull, ts, ws);
                });
                timed(h, K_XSTAGE, [&] { exchange(out[k - 1]); });
            } else {
                // U^n stays in U: the diffusion's boundary faces read it
                timed(h, K_STAGE,
                      [&] { launch_stage<float, float, double, double>(d, t, in[3], U, Un, F, dtp, div, s, pull, ts, ws); });
                timed(h, K_XSTATE, [&] { exchange(F); });
            }
        }
        timed(h, K_DIFF, [&] { launch_diffusion<float, float>(d, t, F, Un, Un, Fn, part, dtp, s, pull, td, wd, U); });
        timed(h, K_XSTATE, [&] { exchange(Fn); });
    }
}

// Bring derived device data in line with U[cur] before stepping.
void prepare_steps(falcon_handle *h) {
    const Dims &d = h->dims;
    const long long n = (long long)d.nblocks * 8 * d.P;
    if (h->precision == 1 && !h->shadow_valid) {
        if (!h->u_ghosts_valid) {
            launch_exchange<double>(d, h->gh64, h->U[h->cur].as<double>(), h->stream);
            h->u_ghosts_valid = true;
        }
        launch_to_f32(h->U[h->cur].as<double>(), h->F32[h->cur].as<float>(), n, h->stream);
        // ghosts as every later FP32 stage input has them: formed by the FP32
        // ghost fill from the rounded interiors, which is also what the
        // kernels' in-place reads give (rounding the FP64 ghosts instead can
        // differ in the last bit where a weight carries rounding residue)
        launch_exchange<float>(d, h->gh32, h->F32[h->cur].as<float>(), h->stream);
        h->shadow_valid = true;
    }
    if (h->precision == 0 && !h->u_ghosts_valid) {
        launch_exchange<double>(d, h->gh64, h->U[h->cur].as<double>(), h->stream);
        h->u_ghosts_valid = true;
    }
    if (!h->partials_valid) {
        launch_dt_partial(d, h->dt, h->U[h->cur].as<double>(), h->partial.as<double>(), h->stream,
                          h->tile_diff[h->precision]);
        h->partials_valid = true;
    }
}

void after_steps(falcon_handle *h) {
    if (h->precision == 1) { h->u_ghosts_valid = false; h->shadow_valid = true; }
    else { h->u_ghosts_valid = true; h->shadow_valid = false; }
    h->partials_valid = true;
}

// After FP32-mode steps the FP64 state's ghosts are refilled from its interior.
void refresh_u_ghosts(falcon_handle *h) {
    if (h->u_ghosts_valid) return;
    double *U = h->U[h->cur].as<double>();
    launch_exchange<double>(h->dims, h->gh64, U, h->stream);
    launch_boundary<double>(h->dims, h->bnd, U, h->stream);
    h->u_ghosts_valid = true;
}

void check_launch() { HIPCHECK(hipGetLastError()); }

}  // namespace

// host Fortran order (i fastest) <-> device [i][j][k] (k fastest) for m^3 blocks
void heat_to_device(const double *src, double *dst, int m, long long nblk) {
    const long long M3 = (long long)m * m * m;
    for (long long b = 0; b < nblk; b++)
        for (int k = 0; k < m; k++)
            for (int j = 0; j < m; j++)
                for (int i = 0; i < m; i++)
                    dst[b * M3 + ((long long)i * m + j) * m + k] = src[b * M3 + i + (long long)m * (j + (long long)m * k)];
}
void heat_to_host(const double *src, double *dst, int m, long long nblk) {
    const long long M3 = (long long)m * m * m;
    for (long long b = 0; b < nblk; b++)
        for (int k = 0; k < m; k++)
            for (int j = 0; j < m; j++)
                for (int i = 0; i < m; i++)
                    dst[b * M3 + i + (long long)m * (j + (long long)m * k)] = src[b * M3 + ((long long)i * m + j) * m + k];
}

extern "C" {

int falcon_last_error(char *buf, int len) {
    if (!buf || len <= 0) return -1;
    std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
    return 0;
}

int falcon_create(falcon_handle_t *out, int ni, int nj, int nk, int ng, double r_lo_cgs, double r_hi_cgs) {
    return guard([&] {
        if (!out) throw Error("null output handle");
        if (ni < 1 || nj < 1 || nk < 1 || ng != 2) throw Error("block sizes must be >= 1 and ng must be 2");
        if (!(r_hi_cgs > r_lo_cgs) || !(r_lo_cgs > 0)) throw Error("invalid radial range");
        auto h = std::make_unique<falcon_handle>();
        h->lay.tree.init(ni, nj, nk, ng, r_lo_cgs, r_hi_cgs);
        *out = h.release();
    });
}

int falcon_destroy(falcon_handle_t h) {
    return guard([&] { delete h; });
}

int falcon_tree_divide_all(falcon_handle_t h, int nr, int nt, int np) {
    return guard([&] {
        if (h->labelled) throw Error("tree already labelled");
        if (nr < 1 || nt < 1 || np < 1 || nr > 64 || nt > 64 || np > 64) throw Error("split counts must be in [1,64]");
        h->lay.tree.divide_all(nr, nt, np);
    });
}

int falcon_tree_amr(falcon_handle_t h, const int flags[3], int if_global, const double rng_rsun[6]) {
    return guard([&] {
        if (h->labelled) throw Error("tree already labelled");
        h->lay.tree.amr(flags, if_global, rng_rsun);
    });
}

int falcon_tree_refine(falcon_handle_t h, int n, const int *leaf_ordinals) {
    return guard([&] {
        if (h->labelled) throw Error("tree already labelled");
        if (n < 0) throw Error("n must be >= 0");
        h->lay.tree.refine(std::vector<int>(leaf_ordinals, leaf_ordinals + n));
    });
}

int falcon_tree_set_nodes(falcon_handle_t h, int n_nodes, const int *parent, const int *is_yin, const double *rng6) {
    return guard([&] {
        if (h->labelled) throw Error("tree already labelled");
        Tree &t = h->lay.tree;
        if (n_nodes < 2) throw Error("need at least the two roots");
        // map given pre-order nodes to the library's node table; children of a
        // node keep their given order.
        std::vector<std::vector<int>> kids(n_nodes);
        std::vector<int> roots;
        for (int n = 0; n < n_nodes; n++) {
            if (parent[n] < 0) roots.push_back(n);
            else {
                if (parent[n] >= n) throw Error("nodes must be in pre-order");
                kids[parent[n]].push_back(n);
            }
        }
        if (roots.size() != 2 || !is_yin[roots[0]] || is_yin[roots[1]]) throw Error("expected Yin root then Yang root");
        t.nodes.clear();
        std::vector<int> map(n_nodes, -1);
        auto mk = [&](int n) {
            Node nd;
            nd.yin = is_yin[n] ? 1 : 0;
            for (int a = 0; a < 3; a++) { nd.rng[a][0] = rng6[6 * n + a]; nd.rng[a][1] = rng6[6 * n + 3 + a]; }
            t.nodes.push_back(nd);
            return (int)t.nodes.size() - 1;
        };
        map[roots[0]] = mk(roots[0]);
        map[roots[1]] = mk(roots[1]);
        // breadth over given nodes: allocate children contiguously
        std::vector<int> queue = {roots[0], roots[1]};
        for (size_t qi = 0; qi < queue.size(); qi++) {
            int n = queue[qi];
            if (kids[n].empty()) continue;
            int first = (int)t.nodes.size();
            for (int c : kids[n]) { map[c] = mk(c); queue.push_back(c); }
            Node &p = t.nodes[map[n]];
            p.leaf = 0;
            p.first_child = first;
            // children per axis: distinct lower bounds along each axis
            int N[3];
            for (int a = 0; a < 3; a++) {
                std::vector<double> lo;
                for (int c : kids[n]) lo.push_back(rng6[6 * c + a]);
                std::sort(lo.begin(), lo.end());
                N[a] = (int)(std::unique(lo.begin(), lo.end()) - lo.begin());
            }
            if (N[0] * N[1] * N[2] != (int)kids[n].size()) throw Error("children do not form a tensor split");
            p.N[0] = N[0]; p.N[1] = N[1]; p.N[2] = N[2];
        }
    });
}

int falcon_nblocks(falcon_handle_t h, int *nblocks, int *nleaf_yin) {
    return guard([&] {
        h->need_label();
        if (nblocks) *nblocks = h->lay.tree.nleaf;
        if (nleaf_yin) *nleaf_yin = h->lay.tree.nleaf_yin;
    });
}

int falcon_block_info(falcon_handle_t h, int ib, double *rng6, int *flags3, double *xi, double *xj, double *xk) {
    return guard([&] {
        h->need_label();
        h->check_block(ib);
        const Tree &t = h->lay.tree;
        const Node &n = t.nodes[t.leaf_node[ib]];
        if (rng6) for (int a = 0; a < 3; a++) { rng6[a] = n.rng[a][0]; rng6[3 + a] = n.rng[a][1]; }
        // centres per ModBlock_InitGrid (same code path as the table builder)
        const int ng = t.ng;
        const int nn[3] = {t.ni, t.nj, t.nk};
        double *xs[3] = {xi, xj, xk};
        for (int a = 0; a < 3; a++) {
            if (!xs[a]) continue;
            for (int i = 1 - ng; i <= nn[a] + ng; i++)
                xs[a][i + ng - 1] = (n.rng[a][0] * (double)((float)(nn[a] - i) + 0.5f) +
                                     n.rng[a][1] * (double)((float)i - 0.5f)) / (double)nn[a];
        }
        if (flags3) {
            double lo = (n.rng[0][0] * (double)((float)(t.ni + ng - 1) + 0.5f) + n.rng[0][1] * (double)((float)(1 - ng) - 0.5f)) / (double)t.ni;
            double hi = (n.rng[0][0] * (double)((float)(-ng) + 0.5f) + n.rng[0][1] * (double)((float)(t.ni + ng) - 0.5f)) / (double)t.ni;
            flags3[0] = n.yin;
            flags3[1] = hi > t.r_range[1];
            flags3[2] = lo < t.r_range[0];
        }
    });
}

int falcon_build_tables(falcon_handle_t h) {
    return guard([&] {
        h->need_label();
        if (h->built) return;
        Layout &L = h->lay;
        L.keep_gcib = L.tree.nleaf <= 4096;
        L.build(host_threads());
        const Tree &t = L.tree;
        h->dims = make_dims(t.ni, t.nj, t.nk, t.ng, t.nleaf);
        h->built = true;
    });
}

int falcon_set_device_tables(falcon_handle_t h, int on) {
    return guard([&] {
        if (h->built) throw Error("the ghost tables are already built");
        h->device_tables = on != 0;
    });
}

int falcon_set_gc_all(falcon_handle_t h) {
    int rc = guard([&] {
        if (!h->stream) {
            HIPCHECK(hipGetDevice(&h->device));
            HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
            scratch_setup(h->device);
        }
        if (h->device_tables && !h->built) {
            h->need_label();
            Layout &L = h->lay;
            L.keep_gcib = L.tree.nleaf <= 4096;
            build_ghosts_device(L, host_threads(), h->stream);
            const Tree &t = L.tree;
            h->dims = make_dims(t.ni, t.nj, t.nk, t.ng, t.nleaf);
            h->built = true;
        }
    });
    if (rc != 0 || falcon_build_tables(h) != 0) return -1;
    return guard([&] {
        alloc_state(h);
        h->tables = true;
    });
}

int falcon_gc_iblocks(falcon_handle_t h, int ib, int32_t *out) {
    return guard([&] {
        h->need_built();
        h->check_block(ib);
        if (h->lay.gcib.empty()) throw Error("ghost donor maps are kept only for layouts of <= 4096 blocks");
        const auto &g = h->lay.gcib[ib];
        std::memcpy(out, g.data(), sizeof(int32_t) * g.size());
    });
}

int falcon_gc_ntargets(falcon_handle_t h, int ib, int *ntargets) {
    return guard([&] {
        h->need_built();
        h->check_block(ib);
        *ntargets = (int)h->lay.targets[ib].size();
    });
}

int falcon_gc_target(falcon_handle_t h, int ib, int it, int *donor, int *ngc, int32_t *ijk) {
    return guard([&] {
        h->need_built();
        h->check_block(ib);
        const auto &T = h->lay.targets[ib];
        if (it < 0 || it >= (int)T.size()) throw Error("target index out of range");
        const int NI = h->NI(), NJ = h->NJ(), ng = h->lay.tree.ng;
        const auto &lst = T[it].second;
        if (donor) *donor = T[it].first;
        if (ngc) *ngc = (int)lst.size();
        if (ijk) {
            const size_t n = lst.size();
            for (size_t g = 0; g < n; g++) {
                int32_t q = lst[g];
                ijk[g] = q % NI - ng + 1;
                ijk[g + n] = (q / NI) % NJ - ng + 1;
                ijk[g + 2 * n] = q / (NI * NJ) - ng + 1;
            }
        }
    });
}

int falcon_ghost_pull_table(falcon_handle_t h, int32_t *pull, int64_t *n_copy, int64_t *n_rest) {
    return guard([&] {
        h->need_built();
        std::vector<int2> copy32, rest;
        std::vector<std::pair<int2, double>> lerp32;
        std::vector<int32_t> tab;
        fp32_copy_lerp(h, copy32, lerp32);
        build_pull_table(h, copy32, tab, rest);
        if (pull) std::copy(tab.begin(), tab.end(), pull);
        if (n_copy) *n_copy = (int64_t)copy32.size();
        if (n_rest) *n_rest = (int64_t)rest.size();
    });
}

int falcon_gc_entries(falcon_handle_t h, int64_t *n, int32_t *recv, int32_t *donor, int32_t *gijk, int32_t *bijk,
                      double *w, int32_t *yy, double *rot) {
    return guard([&] {
        h->need_built();
        const auto &G = h->lay.ghosts;
        *n = (int64_t)G.size();
        if (!recv) return;
        for (size_t e = 0; e < G.size(); e++) {
            const GhostEntry &x = G[e];
            recv[e] = x.recv_block;
            donor[e] = x.donor;
            gijk[3 * e] = x.gi; gijk[3 * e + 1] = x.gj; gijk[3 * e + 2] = x.gk;
            bijk[3 * e] = x.i0; bijk[3 * e + 1] = x.j0; bijk[3 * e + 2] = x.k0;
            w[3 * e] = x.w[0]; w[3 * e + 1] = x.w[1]; w[3 * e + 2] = x.w[2];
            yy[e] = x.yy;
            rot[2 * e] = x.rot[0]; rot[2 * e + 1] = x.rot[1];
        }
    });
}

int falcon_set_background(falcon_handle_t h, int ib, const double *bg) {
    return guard([&] {
        h->need_label();
        h->check_block(ib);
        const int NI = h->NI();
        std::memcpy(h->bg.data() + (size_t)ib * 8 * NI, bg, sizeof(double) * 8 * NI);
        h->bg_set[ib] = 1;
        if (h->dt.meta) h->bg_dirty = true;
    });
}

int falcon_get_background(falcon_handle_t h, int ib, double *bg) {
    return guard([&] {
        h->need_label();
        h->check_block(ib);
        if (!h->bg_set[ib]) throw Error("background not set for block " + std::to_string(ib));
        const int NI = h->NI();
        std::memcpy(bg, h->bg.data() + (size_t)ib * 8 * NI, sizeof(double) * 8 * NI);
    });
}

int falcon_background_polytrope(falcon_handle_t h, double contrast, double rho_top) {
    return guard([&] {
        h->need_label();
        const Tree &t = h->lay.tree;
        PolytropeParams p;
        p.r_lo = t.r_range[0]; p.r_hi = t.r_range[1]; p.contrast = contrast; p.rho_top = rho_top;
        const int NI = h->NI(), ng = t.ng;
        std::vector<double> xi(NI);
        for (int b = 0; b < t.nleaf; b++) {
            const Node &n = t.nodes[t.leaf_node[b]];
            for (int i = 1 - ng; i <= t.ni + ng; i++)
                xi[i + ng - 1] = (n.rng[0][0] * (double)((float)(t.ni - i) + 0.5f) + n.rng[0][1] * (double)((float)i - 0.5f)) / (double)t.ni;
            double *B = h->bg.data() + (size_t)b * 8 * NI;
            for (int I = 0; I < NI; I++) {
                double o[8];
                polytrope(p, xi[I], o);
                for (int a = 0; a < 8; a++) B[(size_t)a * NI + I] = o[a];
            }
            h->bg_set[b] = 1;
        }
        if (h->dt.meta) h->bg_dirty = true;
    });
}

static void ensure_device_tables(falcon_handle_t h) {
    if (!h->dt.meta) upload_tables(h);
    else if (h->bg_dirty) upload_radial(h);
    h->bg_dirty = false;
}

int falcon_set_state(falcon_handle_t h, int ib, const double *prim) {
    return guard([&] {
        h->need_tables();
        h->check_block(ib);
        const size_t P = (size_t)h->dims.P;
        std::vector<double> tmp(8 * P);
        fortran_to_device(prim, tmp.data(), h->dims);
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipMemcpy(h->U[h->cur].as<double>() + (size_t)ib * 8 * P, tmp.data(), sizeof(double) * 8 * P,
                           hipMemcpyHostToDevice));
        h->partials_valid = false;
        h->shadow_valid = false;
    });
}

int falcon_get_state(falcon_handle_t h, int ib, double *prim) {
    return guard([&] {
        h->need_tables();
        h->check_block(ib);
        const size_t P = (size_t)h->dims.P;
        std::vector<double> tmp(8 * P);
        if (h->dt.meta) refresh_u_ghosts(h);
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipMemcpy(tmp.data(), h->U[h->cur].as<double>() + (size_t)ib * 8 * P, sizeof(double) * 8 * P,
                           hipMemcpyDeviceToHost));
        device_to_fortran(tmp.data(), prim, h->dims);
    });
}

int falcon_set_divb(falcon_handle_t h, int option) {
    return guard([&] {
        if (option != 0 && option != 1) throw Error("DivB option must be 0 (none) or 1 (GLM)");
        h->need_tables();
        if (option == 1 && h->precision != 0) throw Error("GLM divergence cleaning runs on the FP64 path only");
        if (option == 1 && !h->psi[0].p) {
            const Dims &d = h->dims;
            const size_t n = (size_t)d.nblocks * d.P, ng = (size_t)d.nblocks * 3 * d.ni * d.nj * d.nk;
            for (DevBuf *b : {&h->psi[0], &h->psi[1], &h->psiS[0], &h->psiS[1]}) {
                b->alloc(n * sizeof(double));
                HIPCHECK(hipMemset(b->p, 0, n * sizeof(double)));
            }
            h->glmG.alloc(ng * sizeof(double));
            HIPCHECK(hipMemset(h->glmG.p, 0, ng * sizeof(double)));
            h->glmC.alloc(ng / 3 * 2 * sizeof(double));
            HIPCHECK(hipMemset(h->glmC.p, 0, ng / 3 * 2 * sizeof(double)));
        }
        h->divb = option;
    });
}

int falcon_set_psi(falcon_handle_t h, int ib, const double *psi) {
    return guard([&] {
        h->need_tables();
        h->check_block(ib);
        if (h->divb != 1) throw Error("psi is stepped only with GLM divergence cleaning (falcon_set_divb(h, 1))");
        const Dims &d = h->dims;
        std::vector<double> tmp((size_t)d.P);
        for (int K = 0; K < d.NK; K++)
            for (int J = 0; J < d.NJ; J++)
                for (int I = 0; I < d.NI; I++)
                    tmp[cell_off(d, I, J, K)] = psi[(size_t)I + (size_t)d.NI * (J + (size_t)d.NJ * K)];
        HIPCHECK(hipStreamSynchronize(h->stream));
        // every psi buffer: the reference starts primitive_rk_IV as a copy of
        // primitive_IV, and no kernel writes psi's radial boundary ghosts
        for (DevBuf *b : {&h->psi[0], &h->psi[1], &h->psiS[0], &h->psiS[1]})
            HIPCHECK(hipMemcpy(b->as<double>() + (size_t)ib * d.P, tmp.data(), sizeof(double) * d.P,
                               hipMemcpyHostToDevice));
    });
}

int falcon_get_psi(falcon_handle_t h, int ib, double *psi) {
    return guard([&] {
        h->need_tables();
        h->check_block(ib);
        if (h->divb != 1) throw Error("psi is stepped only with GLM divergence cleaning (falcon_set_divb(h, 1))");
        const Dims &d = h->dims;
        std::vector<double> tmp((size_t)d.P);
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipMemcpy(tmp.data(), h->psi[h->cur].as<double>() + (size_t)ib * d.P, sizeof(double) * d.P,
                           hipMemcpyDeviceToHost));
        for (int K = 0; K < d.NK; K++)
            for (int J = 0; J < d.NJ; J++)
                for (int I = 0; I < d.NI; I++)
                    psi[(size_t)I + (size_t)d.NI * (J + (size_t)d.NJ * K)] = tmp[cell_off(d, I, J, K)];
    });
}

int falcon_init_synthetic(falcon_handle_t h, uint64_t seed, double s_amp, double b_ratio) {
    return guard([&] {
        h->need_tables();
        ensure_device_tables(h);
        const Dims &d = h->dims;
        const int nb = d.nblocks, NI = h->NI();
        // per block: its ranges, Yin flag and the rows c_p = Gamma3/(Gamma3-1)
        // p0/(rho0 T0), sqrt(rho0), c_s = sqrt(Gamma1 p0/rho0) of its background
        std::vector<double> rng((size_t)6 * nb), rows((size_t)3 * nb * NI), modes;
        std::vector<int32_t> yin(nb);
        for (int b = 0; b < nb; b++) {
            const BlockGeom &g = h->lay.geom[b];
            for (int a = 0; a < 3; a++) { rng[6 * b + 2 * a] = g.rng[a][0]; rng[6 * b + 2 * a + 1] = g.rng[a][1]; }
            yin[b] = g.yin;
            const double *B = h->bg.data() + (size_t)b * 8 * NI;
            double *r = rows.data() + (size_t)3 * b * NI;
            for (int I = 0; I < NI; I++) {
                const double rho0 = B[I + NI], p0 = B[I + 2 * NI], T0 = B[I + 3 * NI];
                const double g1 = B[I + 4 * NI], g3 = B[I + 5 * NI];
                r[I] = g3 / (g3 - 1.0) * p0 / (rho0 * T0);
                r[NI + I] = std::sqrt(rho0);
                r[2 * NI + I] = std::sqrt(g1 * p0 / rho0);
            }
        }
        synthetic_modes(h->lay.tree, seed, modes);
        const int nm = (int)modes.size() / 7;
        std::vector<SynthMode> sm(nm);
        for (int m = 0; m < nm; m++) {
            for (int a = 0; a < 3; a++) { sm[m].k[a] = modes[7 * m + a]; sm[m].a[a] = modes[7 * m + 3 + a]; }
            sm[m].ph = modes[7 * m + 6];
        }
        const long long ncell = (long long)nb * d.ni * d.nj * d.nk;
        DevBuf drng, dyin, drows, dmodes, dmax;
        drng.alloc(sizeof(double) * rng.size());
        dyin.alloc(sizeof(int32_t) * yin.size());
        drows.alloc(sizeof(double) * rows.size());
        dmodes.alloc(sizeof(SynthMode) * nm);
        dmax.alloc(sizeof(double) * (size_t)((ncell + 255) / 256));
        HIPCHECK(hipMemcpy(drng.p, rng.data(), drng.bytes, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dyin.p, yin.data(), dyin.bytes, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(drows.p, rows.data(), drows.bytes, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dmodes.p, sm.data(), dmodes.bytes, hipMemcpyHostToDevice));
        double *U = h->U[h->cur].as<double>();
        HIPCHECK(hipMemsetAsync(U, 0, sizeof(double) * (size_t)nb * 8 * d.P, h->stream));
        launch_synthetic(d, drng.as<double>(), dyin.as<int32_t>(), drows.as<double>(), dmodes.as<SynthMode>(), nm, seed,
                         s_amp, b_ratio, U, dmax.as<double>(), h->stream);
        // the reference driver's two initial fills (3_6/test1.f90:143-144)
        launch_exchange<double>(d, h->gh64, U, h->stream);
        launch_exchange<double>(d, h->gh64, U, h->stream);
        h->partials_valid = false;
        h->shadow_valid = false;
        h->u_ghosts_valid = true;
        check_launch();
        HIPCHECK(hipStreamSynchronize(h->stream));
    });
}

int falcon_set_precision(falcon_handle_t h, int mode) {
    return guard([&] {
        if (mode != 0 && mode != 1) throw Error("precision must be 0 (FP64) or 1 (FP32 deviation)");
        if (mode == 1 && h->divb) throw Error("GLM divergence cleaning runs on the FP64 path only");
        if (mode != h->precision && h->dt.meta) {
            refresh_u_ghosts(h);
            h->shadow_valid = false;
        }
        // the partial minima are laid out per tile of the diffusion's edge
        if (h->tile_diff[mode] != h->tile_diff[h->precision]) h->partials_valid = false;
        h->precision = mode;
    });
}

// Launch shapes of precision p's marches; a change drops that precision's
// graph and, where the diffusion's tile edge changes, the partial minima laid
// out by it.
static void apply_shape(falcon_handle *h, int p, int ts, int ws, int td, int wd) {
    if ((ts != 8 && ts != 16) || (td != 8 && td != 16)) throw Error("tile edges must be 8 or 16");
    if (ws < 0 || wd < 0) throw Error("workgroup counts must be >= 0");
    if (ts == h->tile_stage[p] && td == h->tile_diff[p] && ws == h->wgs_stage[p] && wd == h->wgs_diff[p]) return;
    if (h->graph && h->graph_prec == p) {
        HIPCHECK(hipGraphExecDestroy(h->graph));
        h->graph = nullptr;
    }
    if (p == h->precision && td != h->tile_diff[p]) h->partials_valid = false;
    h->tile_stage[p] = ts;
    h->tile_diff[p] = td;
    h->wgs_stage[p] = ws;
    h->wgs_diff[p] = wd;
}

int falcon_set_tiles(falcon_handle_t h, int precision, int stage_tile, int stage_wgs, int diffusion_tile,
                     int diffusion_wgs) {
    return guard([&] {
        if (precision != 0 && precision != 1) throw Error("precision must be 0 (FP64) or 1 (FP32 deviation)");
        apply_shape(h, precision, stage_tile, stage_wgs, diffusion_tile, diffusion_wgs);
    });
}

int falcon_get_tiles(falcon_handle_t h, int precision, int *shape) {
    return guard([&] {
        if (precision != 0 && precision != 1) throw Error("precision must be 0 (FP64) or 1 (FP32 deviation)");
        if (!shape) throw Error("null output");
        shape[0] = h->tile_stage[precision];
        shape[1] = h->wgs_stage[precision];
        shape[2] = h->tile_diff[precision];
        shape[3] = h->wgs_diff[precision];
    });
}

// Autotuner of the march launch shapes.  For the current precision, times the
// stage kernel (a step's four stage launches) and the diffusion kernel in four
// shapes, tile edge 16 or 8 with one workgroup per tile or with the persistent
// grid (the workgroups the device holds resident at once), reps launches after
// one untimed launch, twice over the shapes, on the current state, and keeps
// the fastest shape of each.  The timed launches write only the stage buffer
// S[0], the next-state buffers and the partial minima, none of which holds anything between steps
// (the partial minima are recomputed before the next step), so the state and
// every later result are unchanged.
// shape_out (optional): stage edge, stage workgroups, diffusion edge,
// diffusion workgroups.  ms_out (optional): ms per stage launch (a step's four
// averaged) in the shapes (16, per tile), (16, persistent), (8, per tile),
// (8, persistent), then ms per diffusion launch in the same order.
int falcon_autotune(falcon_handle_t h, int reps, int *shape_out, double *ms_out) {
    return guard([&] {
        if (reps < 1 || reps > 1000) throw Error("reps must be in 1..1000");
        h->need_tables();
        ensure_device_tables(h);
        alloc_stage(h);
        prepare_steps(h);
        const Dims &d = h->dims;
        const DevTables &t = h->dt;
        hipStream_t s = h->stream;
        const int p = h->precision, cur = h->cur;
        double *U = h->U[cur].as<double>(), *Un = h->U[1 - cur].as<double>();
        double *dtp = h->dtp.as<double>(), *part = h->partial.as<double>();
        const int32_t *pull = (p == 1 && h->ghost_pull) ? h->gh32.pull : nullptr;
        // a time step for the timed launches to scale by; the diffusion launches
        // then overwrite the partial minima
        launch_dt_final(d, part, dtp, 0.5, s, h->tile_diff[p]);
        h->partials_valid = false;
        // destroyed on every exit, an error included
        struct Ev {
            hipEvent_t e = nullptr;
            Ev() { HIPCHECK(hipEventCreate(&e)); }
            ~Ev() { if (e) (void)hipEventDestroy(e); }
        } ev0, ev1;
        hipEvent_t e0 = ev0.e, e1 = ev1.e;
        auto time = [&](auto &&launch) -> double {
            launch();
            HIPCHECK(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) launch();
            HIPCHECK(hipEventRecord(e1, s));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / reps;
        };
        // shape c: edge 16 (c < 2) or 8, one workgroup per tile (c even) or persistent
        int edge[4], wst[4], wdi[4];
        double ms[8];
        for (int c = 0; c < 4; c++) {
            edge[c] = c < 2 ? 16 : 8;
            wst[c] = (c & 1) ? march_resident_wgs(0, p, edge[c]) : 0;
            wdi[c] = (c & 1) ? march_resident_wgs(1, p, edge[c]) : 0;
        }
        // the stage is timed as a step runs it: three intermediate stages and the
        // final one (for the FP32 path its FP64 form, writing the next-state
        // buffers); ms is per stage launch.  Two sweeps over the shapes, the
        // faster time of each kept, so the order of the shapes does not bias them.
        for (int q = 0; q < 8; q++) ms[q] = 1e30;
        for (int sweep = 0; sweep < 2; sweep++) {
            for (int c = 0; c < 4; c++) {
                const int tl = edge[c], ws = wst[c], wd = wdi[c];
                double a, b;
                if (p == 0) {
                    double *S0 = h->S64[0].as<double>();
                    a = time([&] {
                        for (int k = 1; k <= 3; k++)
                            launch_stage<double, double, double, double>(d, t, U, U, S0, nullptr, dtp, 5.0 - k, s,
                                                                         nullptr, tl, ws);
                        launch_stage<double, double, double, double>(d, t, U, U, Un, nullptr, dtp, 1.0, s, nullptr, tl,
                                                                     ws);
                    }) / 4;
                    b = time([&] {
                        launch_diffusion<double, double>(d, t, U, U, Un, nullptr, part, dtp, s, nullptr, tl, wd);
                    });
                } else {
                    float *F = h->F32[cur].as<float>(), *Fn = h->F32[1 - cur].as<float>(), *S0 = h->S32[0].as<float>();
                    a = time([&] {
                        for (int k = 1; k <= 3; k++)
                            launch_stage<float, float, float, float>(d, t, F, F, S0, nullptr, dtp, 5.0 - k, s, pull, tl,
                                                                     ws);
                        launch_stage<float, float, double, double>(d, t, F, U, Un, Fn, dtp, 1.0, s, pull, tl, ws);
                    }) / 4;
                    b = time([&] {
                        launch_diffusion<float, float>(d, t, F, Un, Un, Fn, part, dtp, s, pull, tl, wd, U);
                    });
                }
                ms[c] = std::min(ms[c], a);
                ms[4 + c] = std::min(ms[4 + c], b);
            }
        }
        check_launch();
        int bs = 0, bd = 0;
        for (int c = 1; c < 4; c++) {
            if (ms[c] < ms[bs]) bs = c;
            if (ms[4 + c] < ms[4 + bd]) bd = c;
        }
        apply_shape(h, p, edge[bs], wst[bs], edge[bd], wdi[bd]);
        if (shape_out) {
            shape_out[0] = edge[bs]; shape_out[1] = wst[bs];
            shape_out[2] = edge[bd]; shape_out[3] = wdi[bd];
        }
        if (ms_out) for (int q = 0; q < 8; q++) ms_out[q] = ms[q];
    });
}

int falcon_set_ghost_pull(falcon_handle_t h, int on) {
    return guard([&] {
        if ((on != 0) != h->ghost_pull && h->graph) {
            HIPCHECK(hipGraphExecDestroy(h->graph));
            h->graph = nullptr;
        }
        if (on && !h->ghost_pull) {
            // With the pull on, the ghosts that the kernels read in place are
            // never written again, yet the FP32 Yin-Yang gather reads some of
            // them as zero-weight corners (f * 0).  Whatever a run without the
            // pull left there (a diverged run's NaN included) is cleared: the
            // FP32 buffers restart from zeros and the shadow of U is rebuilt.
            if (h->stream) HIPCHECK(hipStreamSynchronize(h->stream));
            for (DevBuf *b : {&h->S32[0], &h->S32[1], &h->F32[0], &h->F32[1]})
                if (b->p) HIPCHECK(hipMemset(b->p, 0, b->bytes));
            h->shadow_valid = false;
        }
        h->ghost_pull = on != 0;
    });
}

int falcon_sendrecv_gc(falcon_handle_t h, int which) {
    return guard([&] {
        h->need_tables();
        ensure_device_tables(h);
        if (which == 0) {
            if (h->divb)
                launch_exchange_glm(h->dims, h->gh64, h->U[h->cur].as<double>(), h->psi[h->cur].as<double>(), h->stream);
            else launch_exchange<double>(h->dims, h->gh64, h->U[h->cur].as<double>(), h->stream);
            h->shadow_valid = false;
        } else {
            alloc_stage(h);
            if (h->precision == 0 && h->divb)
                launch_exchange_glm(h->dims, h->gh64, h->S64[0].as<double>(), h->psiS[0].as<double>(), h->stream);
            else if (h->precision == 0) launch_exchange<double>(h->dims, h->gh64, h->S64[0].as<double>(), h->stream);
            else launch_exchange<float>(h->dims, h->gh32, h->S32[0].as<float>(), h->stream);
        }
        check_launch();
        HIPCHECK(hipStreamSynchronize(h->stream));
    });
}

int falcon_compute_dt(falcon_handle_t h, double cfl, double *dt) {
    return guard([&] {
        h->need_tables();
        ensure_device_tables(h);
        const int td = h->tile_diff[h->precision];
        launch_dt_partial(h->dims, h->dt, h->U[h->cur].as<double>(), h->partial.as<double>(), h->stream, td);
        launch_dt_final(h->dims, h->partial.as<double>(), h->dtp.as<double>(), cfl, h->stream, td);
        h->partials_valid = true;
        check_launch();
        HIPCHECK(hipMemcpyAsync(dt, h->dtp.p, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
    });
}

int falcon_advance(falcon_handle_t h, int nsteps, double cfl, int use_graph, double *dt_last) {
    return guard([&] {
        h->need_tables();
        ensure_device_tables(h);
        alloc_stage(h);
        if (nsteps < 0) throw Error("nsteps < 0");
        if (nsteps == 0) return;
        prepare_steps(h);
        if (use_graph && nsteps >= 2) {
            if (!h->graph || h->graph_cur != h->cur || h->graph_prec != h->precision || h->graph_cfl != cfl ||
                h->graph_fixed_dt != h->fixed_dt || h->graph_divb != h->divb) {
                if (h->graph) { HIPCHECK(hipGraphExecDestroy(h->graph)); h->graph = nullptr; }
                hipGraph_t g;
                HIPCHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                enqueue_step(h, cfl, h->cur);
                enqueue_step(h, cfl, 1 - h->cur);
                HIPCHECK(hipStreamEndCapture(h->stream, &g));
                HIPCHECK(hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0));
                HIPCHECK(hipGraphDestroy(g));
                h->graph_cur = h->cur; h->graph_prec = h->precision; h->graph_cfl = cfl; h->graph_divb = h->divb;
                h->graph_fixed_dt = h->fixed_dt;
            }
            int pairs = nsteps / 2;
            for (int p = 0; p < pairs; p++) HIPCHECK(hipGraphLaunch(h->graph, h->stream));
            for (int s = 2 * pairs; s < nsteps; s++) { enqueue_step(h, cfl, h->cur); h->cur = 1 - h->cur; }
        } else {
            for (int s = 0; s < nsteps; s++) { enqueue_step(h, cfl, h->cur); h->cur = 1 - h->cur; }
        }
        after_steps(h);
        check_launch();
        if (dt_last) {
            HIPCHECK(hipMemcpyAsync(dt_last, h->dtp.p, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        }
        HIPCHECK(hipStreamSynchronize(h->stream));
    });
}

int falcon_compute_dt_noalloc(falcon_handle_t h, double cfl, double *dt) {
    return guard([&] {
        h->need_tables();
        if (!dt) throw Error("falcon_compute_dt_noalloc: null dt");
        if (!h->dt.meta)
            throw Error("falcon_compute_dt_noalloc: the device tables are not uploaded yet (a first "
                        "falcon_compute_dt, falcon_sendrecv_gc or step uploads them)");
        if (h->bg_dirty) {   // rewrites the existing radial rows in place, allocates nothing
            upload_radial(h);
            h->bg_dirty = false;
        }
        int dev = -1;
        HIPCHECK(hipGetDevice(&dev));
        if (dev != h->device) throw Error("falcon_compute_dt_noalloc: call it with the handle's GPU current");
        double *base = g_scratch_ptr[h->device];
        if (!base) throw Error("falcon_compute_dt_noalloc: no dt scratch on this GPU");
        const Dims &d = h->dims;
        const int td = h->tile_diff[h->precision];
        const long long n = (long long)d.nblocks * ((d.nj + td - 1) / td) * ((d.nk + td - 1) / td);
        if (n > DT_NOALLOC_MAX_PARTIALS)
            throw Error("falcon_compute_dt_noalloc: " + std::to_string(n) + " partial minima exceed its limit of " +
                        std::to_string(DT_NOALLOC_MAX_PARTIALS) + " (falcon_compute_dt has none)");
        // every partial is written by the tile pass before the final pass reads it
        double *part = base, *dtp = base + DT_NOALLOC_MAX_PARTIALS;
        launch_dt_partial(d, h->dt, h->U[h->cur].as<double>(), part, h->stream, td);
        launch_dt_final(d, part, dtp, cfl, h->stream, td);
        check_launch();
        HIPCHECK(hipMemcpyAsync(dt, dtp, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
    });
}

int falcon_check_state(falcon_handle_t h, int which, int64_t *n_nan, int32_t *where) {
    return guard([&] {
        h->need_tables();
        if (which != 0 && which != 1) throw Error("falcon_check_state: which must be 0 (state) or 1 (stage buffer)");
        if (!n_nan) throw Error("falcon_check_state: null count");
        if (!h->scanbuf.p) h->scanbuf.alloc(2 * sizeof(unsigned long long));
        unsigned long long init[2] = {0ull, ~0ull}, out[2];
        unsigned long long *w = h->scanbuf.as<unsigned long long>();
        const Dims &d = h->dims;
        if (which == 0 && h->dt.meta) refresh_u_ghosts(h);
        HIPCHECK(hipMemcpyAsync(w, init, sizeof(init), hipMemcpyHostToDevice, h->stream));
        if (which == 0) {
            launch_nan_scan<double>(d, h->U[h->cur].as<double>(), w, h->stream);
            if (h->divb) launch_nan_scan_psi(d, h->psi[h->cur].as<double>(), w, h->stream);
        } else if (h->precision == 0) {
            if (!h->S64[0].p) throw Error("falcon_check_state: no stage buffer yet");
            launch_nan_scan<double>(d, h->S64[0].as<double>(), w, h->stream);
            if (h->divb) launch_nan_scan_psi(d, h->psiS[0].as<double>(), w, h->stream);
        } else {
            if (!h->S32[0].p) throw Error("falcon_check_state: no stage buffer yet");
            launch_nan_scan<float>(d, h->S32[0].as<float>(), w, h->stream);
        }
        check_launch();
        HIPCHECK(hipMemcpyAsync(out, w, sizeof(out), hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
        *n_nan = (int64_t)out[0];
        if (where) {
            for (int q = 0; q < 5; q++) where[q] = -1;
            if (out[0]) {
                unsigned long long k = out[1];
                where[4] = (int32_t)(k % 9); k /= 9;
                where[1] = (int32_t)(k % d.NI); k /= d.NI;
                where[2] = (int32_t)(k % d.NJ); k /= d.NJ;
                where[3] = (int32_t)(k % d.NK); k /= d.NK;
                where[0] = (int32_t)k;
            }
        }
    });
}

int falcon_save_allcells(falcon_handle_t h, int istep, const char *outdir, int merged) {
    return guard([&] {
        h->need_tables();
        if (!h->dt.meta) throw Error("falcon_save_allcells: call falcon_set_gc_all and set a state first");
        if (istep < 0 || istep > 99999999) throw Error("falcon_save_allcells: step must be in [0, 99999999]");
        if (!outdir) throw Error("falcon_save_allcells: null directory");
        ensure_device_tables(h);
        refresh_u_ghosts(h);
        const Dims &d = h->dims;
        const int nb = d.nblocks;
        AllCellsSpec spec{d.ni, d.nj, d.nk, 12, nb,
                          "VARIABLES = \"X [Rs]\", \"Y [Rs]\", \"Z [Rs]\", \"Rho1\", \"Vx\", \"Vy\", \"Vz\", "
                          "\"Bx\", \"By\", \"Bz\", \"S1\", \"PHI\""};
        std::vector<double> geo;
        std::vector<int32_t> yin;
        allcells_geometry(h->lay, geo, yin);
        DevBuf dgeo, dyin, dout;
        dgeo.alloc(sizeof(double) * geo.size());
        dyin.alloc(sizeof(int32_t) * yin.size());
        HIPCHECK(hipMemcpy(dgeo.p, geo.data(), dgeo.bytes, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dyin.p, yin.data(), dyin.bytes, hipMemcpyHostToDevice));
        const size_t per = (size_t)(d.ni + 1) * (d.nj + 1) * (d.nk + 1) * 12;
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nb, ((size_t)256 << 20) / (per * sizeof(double))));
        dout.alloc(sizeof(double) * per * chunk);
        std::vector<double> host(per * chunk);
        char stepname[16];
        std::snprintf(stepname, sizeof(stepname), "%08d", istep);
        const std::filesystem::path dir = std::filesystem::path(outdir) / (std::string("AllCells_") + stepname);
        std::filesystem::create_directories(dir);
        auto put = [&](const std::filesystem::path &p, const std::string &text, bool append) {
            std::ofstream f(p, append ? std::ios::binary | std::ios::app : std::ios::binary | std::ios::trunc);
            f.write(text.data(), (std::streamsize)text.size());
            if (!f) throw Error("falcon_save_allcells: cannot write " + p.string());
        };
        put(dir / "info.dat", allcells_info(spec), false);
        const std::filesystem::path mpath = dir / "merged.dat";
        if (merged) {
            const long long ncorner = (long long)nb * (d.ni + 1) * (d.nj + 1) * (d.nk + 1);
            const long long ncell = (long long)nb * d.ni * d.nj * d.nk;
            put(mpath, "TITLE = \"AllCells\"\n" + spec.variables + "\nZONE N=" + std::to_string(ncorner) +
                           ", E=" + std::to_string(ncell) + ", F=FEPOINT, ET=BRICK\n", false);
        }
        const int nt = std::max(1, std::min(host_threads(), 16));
        for (int b0 = 0; b0 < nb; b0 += chunk) {
            const int n = std::min(chunk, nb - b0);
            launch_allcells(d, h->U[h->cur].as<double>(), dgeo.as<double>(), dyin.as<int32_t>(), b0, n,
                            (double)6.955e10f, dout.as<double>(), h->stream);
            check_launch();
            HIPCHECK(hipMemcpyAsync(host.data(), dout.p, sizeof(double) * per * n, hipMemcpyDeviceToHost, h->stream));
            HIPCHECK(hipStreamSynchronize(h->stream));
            std::vector<std::string> text(n);
            std::vector<std::thread> th;
            for (int w = 0; w < std::min(nt, n); w++)
                th.emplace_back([&, w]() {
                    for (int q = w; q < n; q += nt) allcells_block_text(spec, host.data() + per * q, text[q]);
                });
            for (auto &x : th) x.join();
            for (int q = 0; q < n; q++) {
                if (merged) {
                    put(mpath, text[q], true);
                } else {
                    char name[32];
                    std::snprintf(name, sizeof(name), "Block_%08d.dat", b0 + q + 1);
                    put(dir / name, text[q], false);
                }
            }
        }
        if (merged) {
            for (int b0 = 0; b0 < nb; b0 += 256) {
                std::string conn;
                allcells_connectivity(spec, b0, std::min(nb, b0 + 256), conn);
                put(mpath, conn, true);
            }
            put(dir / ".merged", "", false);
        }
    });
}

int falcon_format_es16(const double *x, int n, char *out) {
    return guard([&] {
        if (n < 0 || (n > 0 && (!x || !out))) throw Error("falcon_format_es16: bad arguments");
        for (int q = 0; q < n; q++) format_es16(x[q], out + 16 * q);
        out[16 * n] = '\0';
    });
}

int falcon_set_fixed_dt(falcon_handle_t h, double dt) {
    return guard([&] {
        if (!(dt >= 0) || !std::isfinite(dt)) throw Error("fixed dt must be finite and >= 0");
        h->fixed_dt = dt;
    });
}

int falcon_state_diff(falcon_handle_t a, falcon_handle_t b, double rel_tol, double *max_diff, double *max_ref,
                      int64_t *n_above, int64_t *argmax) {
    return guard([&] {
        a->need_tables();
        b->need_tables();
        const Dims &d = a->dims, &e = b->dims;
        if (d.ni != e.ni || d.nj != e.nj || d.nk != e.nk || d.ng != e.ng || d.nblocks != e.nblocks)
            throw Error("falcon_state_diff: the two handles have different layouts");
        HIPCHECK(hipStreamSynchronize(b->stream));
        if (!a->diffbuf.p) a->diffbuf.alloc(32 * sizeof(unsigned long long));
        unsigned long long init[32];
        for (int q = 0; q < 32; q++) init[q] = q < 24 ? 0ull : ~0ull;
        unsigned long long *w = a->diffbuf.as<unsigned long long>();
        HIPCHECK(hipMemcpyAsync(w, init, sizeof(init), hipMemcpyHostToDevice, a->stream));
        const double *pa = a->U[a->cur].as<double>(), *pb = b->U[b->cur].as<double>();
        launch_state_diff(d, pa, pb, rel_tol, w, 0, a->stream);
        launch_state_diff(d, pa, pb, rel_tol, w, 1, a->stream);
        check_launch();
        unsigned long long out[32];
        HIPCHECK(hipMemcpyAsync(out, w, sizeof(out), hipMemcpyDeviceToHost, a->stream));
        HIPCHECK(hipStreamSynchronize(a->stream));
        for (int v = 0; v < 8; v++) {
            double x, y;
            std::memcpy(&x, &out[v], 8);
            std::memcpy(&y, &out[8 + v], 8);
            if (max_diff) max_diff[v] = x;
            if (max_ref) max_ref[v] = y;
            if (n_above) n_above[v] = (int64_t)out[16 + v];
            if (argmax) argmax[v] = out[24 + v] == ~0ull ? -1 : (int64_t)out[24 + v];
        }
    });
}

int falcon_advance_rk4(falcon_handle_t h, double cfl, double *dt_out) {
    return falcon_advance(h, 1, cfl, 0, dt_out);
}

int falcon_rhs(falcon_handle_t h, double *r_out) {
    return guard([&] {
        h->need_tables();
        ensure_device_tables(h);
        const Dims &d = h->dims;
        const size_t nn = (size_t)d.ni * d.nj * d.nk;
        h->rbuf.alloc(sizeof(double) * nn * 8 * d.nblocks);
        refresh_u_ghosts(h);
        double *U = h->U[h->cur].as<double>();
        launch_boundary<double>(d, h->bnd, U, h->stream);
        const int ts = h->tile_stage[h->precision], ws = h->wgs_stage[h->precision];
        if (h->precision == 0) launch_rhs_only<double, double>(d, h->dt, U, h->rbuf.as<double>(), h->stream, ts, ws);
        else launch_rhs_only<float, double>(d, h->dt, U, h->rbuf.as<double>(), h->stream, ts, ws);
        check_launch();
        std::vector<double> tmp(nn * 8 * d.nblocks);
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipMemcpy(tmp.data(), h->rbuf.p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
        // device [b][v][i][j][k] -> Fortran (ni,nj,nk,8) per block
        for (int b = 0; b < d.nblocks; b++)
            for (int v = 0; v < 8; v++)
                for (int i = 0; i < d.ni; i++)
                    for (int j = 0; j < d.nj; j++)
                        for (int k = 0; k < d.nk; k++)
                            r_out[((size_t)b * 8 + v) * nn + (size_t)i + (size_t)d.ni * (j + (size_t)d.nj * k)] =
                                tmp[((size_t)b * 8 + v) * nn + ((size_t)i * d.nj + j) * d.nk + k];
    });
}

int falcon_diffusion_rhs(falcon_handle_t h, double *r_out) {
    return guard([&] {
        h->need_tables();
        ensure_device_tables(h);
        const Dims &d = h->dims;
        const size_t nn = (size_t)d.ni * d.nj * d.nk;
        h->rbuf.alloc(sizeof(double) * nn * 8 * d.nblocks);
        double *U = h->U[h->cur].as<double>();
        const int td = h->tile_diff[h->precision], wd = h->wgs_diff[h->precision];
        if (h->precision == 0) launch_diffusion_only<double>(d, h->dt, U, h->rbuf.as<double>(), h->stream, td, wd);
        else launch_diffusion_only<float>(d, h->dt, U, h->rbuf.as<double>(), h->stream, td, wd);
        check_launch();
        std::vector<double> tmp(nn * 8 * d.nblocks);
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipMemcpy(tmp.data(), h->rbuf.p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
        for (int b = 0; b < d.nblocks; b++)
            for (int v = 0; v < 8; v++)
                for (int i = 0; i < d.ni; i++)
                    for (int j = 0; j < d.nj; j++)
                        for (int k = 0; k < d.nk; k++)
                            r_out[((size_t)b * 8 + v) * nn + (size_t)i + (size_t)d.ni * (j + (size_t)d.nj * k)] =
                                tmp[((size_t)b * 8 + v) * nn + ((size_t)i * d.nj + j) * d.nk + k];
    });
}

int falcon_profile_step(falcon_handle_t h, int nsteps, double cfl, double *out_ms) {
    return guard([&] {
        h->need_tables();
        ensure_device_tables(h);
        alloc_stage(h);
        prepare_steps(h);
        double acc[K_NKIND] = {0}, cnt[K_NKIND] = {0}, whole = 0;
        for (int st = 0; st < nsteps; st++) {
            std::vector<hipEvent_t> ev;
            std::vector<int> kind;
            h->prof_ev = &ev;
            h->prof_kind = &kind;
            hipEvent_t a, b;
            HIPCHECK(hipEventCreate(&a));
            HIPCHECK(hipEventCreate(&b));
            HIPCHECK(hipEventRecord(a, h->stream));
            try {
                enqueue_step(h, cfl, h->cur);
            } catch (...) {
                h->prof_ev = nullptr; h->prof_kind = nullptr;
                throw;
            }
            h->prof_ev = nullptr; h->prof_kind = nullptr;
            HIPCHECK(hipEventRecord(b, h->stream));
            h->cur = 1 - h->cur;
            HIPCHECK(hipStreamSynchronize(h->stream));
            float ms;
            for (size_t q = 0; q < kind.size(); q++) {
                HIPCHECK(hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]));
                acc[kind[q]] += ms;
                cnt[kind[q]] += 1;
            }
            HIPCHECK(hipEventElapsedTime(&ms, a, b));
            whole += ms;
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
        after_steps(h);
        const int order[6] = {K_DT, K_BC, K_STAGE, K_XSTAGE, K_XSTATE, K_DIFF};
        for (int q = 0; q < 6; q++) out_ms[q] = cnt[order[q]] > 0 ? acc[order[q]] / cnt[order[q]] : 0.0;
        out_ms[6] = nsteps > 0 ? whole / nsteps : 0.0;
    });
}

int falcon_device_bytes(falcon_handle_t h, int64_t *bytes) {
    return guard([&] {
        const DevBuf *all[] = {&h->meta, &h->radial, &h->angular, &h->ghost64[0], &h->ghost64[1], &h->ghost64[2],
                               &h->ghost64[3], &h->ghost32[0], &h->ghost32[1], &h->ghost32[2], &h->ghost32[3],
                               &h->ghost_copy32, &h->ghost_lerp32[0], &h->ghost_lerp32[1],
                               &h->ghost_rest32, &h->ghost_pull32,
                               &h->bnd_list, &h->U[0], &h->U[1], &h->F32[0], &h->F32[1], &h->S64[0], &h->S64[1],
                               &h->S32[0], &h->S32[1], &h->dtp, &h->partial, &h->rbuf};
        int64_t s = 0;
        for (const DevBuf *b : all) s += (int64_t)b->bytes;
        *bytes = s;
    });
}

// ---------------------------------------------------------------- heat conduction

int falcon_heat_setup(falcon_handle_t h, double kappa0, double r_ref_rsun, int nsys) {
    if (falcon_build_tables(h) != 0) return -1;
    return guard([&] {
        if (nsys < 1) throw Error("nsys must be >= 1");
        if (!(kappa0 > 0)) throw Error("kappa0 must be positive");
        h->heat_params = HeatParams{};
        h->heat_params.kappa0 = kappa0;
        h->heat_params.r_ref = r_ref_rsun * (double)6.955e10f;
        h->heat.reset(new HeatSolver());
        h->heat->setup(h->lay, h->heat_params, nsys, host_threads());
        h->pfss_on = false;
    });
}

// ---------------------------------------------------------------- PFSS potential field

int falcon_pfss_setup(falcon_handle_t h) {
    if (falcon_build_tables(h) != 0) return -1;
    return guard([&] {
        HeatParams hp;
        hp.kappa0 = 1.0;
        hp.r_ref = 1.0;
        hp.kappa_power = 0;
        hp.top_dirichlet = true;
        h->heat_params = hp;
        h->heat.reset(new HeatSolver());
        h->heat->setup(h->lay, hp, 1, host_threads());
        h->heat->set_dt({INFINITY});   // 1/dt = 0: no V/dt term
        build_pfss_tables(h->lay, h->pfss);
        h->pfss_br.assign(h->pfss.bottom.size() * (size_t)h->pfss.m * h->pfss.m, 0.0);
        h->heat->set_rhs(std::vector<double>(h->heat->cells(), 0.0));
        h->pfss_on = true;
    });
}

static HeatSolver &need_pfss(falcon_handle_t h) {
    if (!h->heat || !h->pfss_on) throw Error("falcon_pfss_setup has not been called");
    return *h->heat;
}

int falcon_pfss_bottom_blocks(falcon_handle_t h, int32_t *blocks, int *n) {
    return guard([&] {
        need_pfss(h);
        if (n) *n = (int)h->pfss.bottom.size();
        if (blocks) std::copy(h->pfss.bottom.begin(), h->pfss.bottom.end(), blocks);
    });
}

int falcon_pfss_set_magnetogram(falcon_handle_t h, const double *br) {
    return guard([&] {
        HeatSolver &H = need_pfss(h);
        const PfssTables &t = h->pfss;
        const int m = t.m;
        const long long mm = (long long)m * m;
        h->pfss_br.assign(br, br + t.bottom.size() * mm);
        // b = Br S on the first radial cell of each bottom block (3_5/ModPFSS.f90:43-44)
        std::vector<double> b(H.cells(), 0.0);
        for (size_t q = 0; q < t.bottom.size(); q++)
            for (int k = 0; k < m; k++)
                for (int j = 0; j < m; j++) {
                    const long long f = (long long)q * mm + (long long)k * m + j;
                    b[(long long)t.bottom[q] * mm * m + (long long)j * m + k] = br[f] * t.area[f];
                }
        H.set_rhs(b);
    });
}

int falcon_pfss_bottom_areas(falcon_handle_t h, double *area) {
    return guard([&] {
        need_pfss(h);
        std::copy(h->pfss.area.begin(), h->pfss.area.end(), area);
    });
}

int falcon_pfss_dipole_magnetogram(falcon_handle_t h, int n, const double *xyz, const double *q, const double *l,
                                   double *br) {
    return guard([&] {
        need_pfss(h);
        if (n < 0) throw Error("the number of dipoles must be >= 0");
        std::vector<double> out;
        dipole_magnetogram(h->lay, h->pfss, n, xyz, q, l, out);
        std::copy(out.begin(), out.end(), br);
    });
}

int falcon_pfss_solve(falcon_handle_t h, double tol, int max_iter, int32_t *iters, double *relres, double *seconds) {
    return guard([&] {
        HeatSolver &H = need_pfss(h);
        HeatResult r = H.solve(tol, max_iter, false);
        if (iters) *iters = r.iters[0];
        if (relres) *relres = r.relres[0];
        if (seconds) *seconds = r.seconds;
    });
}

int falcon_pfss_get_phi(falcon_handle_t h, double *phi) {
    return guard([&] {
        HeatSolver &H = need_pfss(h);
        std::vector<double> dev;
        H.get_solution(0, dev);
        heat_to_host(dev.data(), phi, H.m(), (long long)dev.size() / ((long long)H.m() * H.m() * H.m()));
    });
}

int falcon_pfss_field(falcon_handle_t h, double *B) {
    return guard([&] {
        HeatSolver &H = need_pfss(h);
        const PfssTables &t = h->pfss;
        const int m = t.m, nb = h->lay.tree.nleaf;
        const long long mm = (long long)m * m;
        std::vector<double> bot((size_t)nb * mm, 0.0), dev;
        for (size_t q = 0; q < t.bottom.size(); q++)
            for (int k = 0; k < m; k++)
                for (int j = 0; j < m; j++)
                    bot[(long long)t.bottom[q] * mm + (long long)j * m + k] =
                        h->pfss_br[(long long)q * mm + (long long)k * m + j] * t.dr[t.bottom[q]];
        H.pfss_field(bot, t.geo, dev);
        heat_to_host(dev.data(), B, m, 3LL * nb);
    });
}

// any: the operator-level calls (info, apply, V-cycle, graph and coarse-solve
// settings) also act on the solver falcon_pfss_setup made; the heat-only calls
// (dt, T, solve, training) refuse it
static HeatSolver &need_heat(falcon_handle_t h, bool any = false) {
    if (!h->heat) throw Error("falcon_heat_setup has not been called");
    if (h->pfss_on && !any)
        throw Error("the handle's solver is set up for PFSS (falcon_pfss_setup); call falcon_heat_setup first");
    return *h->heat;
}

int falcon_heat_info(falcon_handle_t h, int *levels, int64_t *cells, double *dt_explicit) {
    return guard([&] {
        HeatSolver &H = need_heat(h, true);
        if (levels) *levels = H.levels();
        if (cells) *cells = H.cells();
        if (dt_explicit) *dt_explicit = H.explicit_dt();
    });
}

int falcon_heat_set_dt(falcon_handle_t h, const double *dt) {
    return guard([&] {
        HeatSolver &H = need_heat(h);
        H.set_dt(std::vector<double>(dt, dt + H.nsys()));
    });
}

int falcon_heat_set_T(falcon_handle_t h, const double *T) {
    return guard([&] {
        HeatSolver &H = need_heat(h);
        const long long n = H.cells() * H.nsys();
        std::vector<double> dev(n);
        heat_to_device(T, dev.data(), H.m(), n / ((long long)H.m() * H.m() * H.m()));
        H.set_rhs_from_T(dev);
    });
}

int falcon_heat_seed_T(falcon_handle_t h, uint64_t seed, double sigma) {
    return guard([&] {
        HeatSolver &H = need_heat(h);
        const long long n = H.cells() * H.nsys();
        std::vector<double> T(n);
        // Gaussian T' (Box-Muller on splitmix64 draws), one stream per cell
        for (long long c = 0; c < n; c++) {
            uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(c + 1);
            auto mix = [](uint64_t x) {
                x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
                x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
                return x ^ (x >> 31);
            };
            const double u1 = ((mix(z) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
            const double u2 = (mix(z ^ 0xD1B54A32D192ED03ull) >> 11) * (1.0 / 9007199254740992.0);
            T[c] = sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        }
        H.set_rhs_from_T(T);
    });
}

int falcon_heat_solve(falcon_handle_t h, double tol, int max_iter, int use_corrector, int32_t *iters, double *relres,
                      double *seconds) {
    return guard([&] {
        HeatSolver &H = need_heat(h);
        HeatResult r = H.solve(tol, max_iter, use_corrector != 0);
        for (int s = 0; s < H.nsys(); s++) {
            if (iters) iters[s] = r.iters[s];
            if (relres) relres[s] = r.relres[s];
        }
        if (seconds) *seconds = r.seconds;
    });
}

int falcon_heat_get_T(falcon_handle_t h, int sys, double *T) {
    return guard([&] {
        HeatSolver &H = need_heat(h);
        if (sys < 0 || sys >= H.nsys()) throw Error("system index out of range");
        std::vector<double> dev;
        H.get_solution(sys, dev);
        heat_to_host(dev.data(), T, H.m(), (long long)dev.size() / ((long long)H.m() * H.m() * H.m()));
    });
}

int falcon_heat_apply(falcon_handle_t h, const double *x, double *y) {
    return guard([&] {
        HeatSolver &H = need_heat(h, true);
        const long long n = H.cells() * H.nsys();
        const long long nblk = n / ((long long)H.m() * H.m() * H.m());
        std::vector<double> xd(n), yd;
        heat_to_device(x, xd.data(), H.m(), nblk);
        H.apply(xd, yd);
        heat_to_host(yd.data(), y, H.m(), nblk);
    });
}

int falcon_heat_vcycle(falcon_handle_t h, const double *r, double *z) {
    return guard([&] {
        HeatSolver &H = need_heat(h, true);
        const long long n = H.cells() * H.nsys();
        const long long nblk = n / ((long long)H.m() * H.m() * H.m());
        std::vector<double> rd(n), zd;
        heat_to_device(r, rd.data(), H.m(), nblk);
        H.vcycle_host(rd, zd);
        heat_to_host(zd.data(), z, H.m(), nblk);
    });
}

// ---------------------------------------------------------------- learned correction

static CnnModel &need_cnn(falcon_handle_t h) {
    if (!h->cnn) h->cnn.reset(new CnnModel());
    return *h->cnn;
}

// device copy of a host array, freed on scope exit
struct DevTmp {
    void *p = nullptr;
    explicit DevTmp(size_t bytes) { HIPCHECK(hipMalloc(&p, std::max<size_t>(bytes, 1))); }
    ~DevTmp() { if (p) (void)hipFree(p); }
};

int falcon_cnn_init(falcon_handle_t h, uint64_t seed) {
    return guard([&] { need_cnn(h).init(seed); });
}

int falcon_cnn_set_params(falcon_handle_t h, const float *params) {
    return guard([&] { need_cnn(h).set_params(params); });
}

int falcon_cnn_get_params(falcon_handle_t h, float *params) {
    return guard([&] { need_cnn(h).get_params(params); });
}

int falcon_cnn_forward(falcon_handle_t h, int64_t nimg, const float *x, float *out) {
    return guard([&] {
        if (nimg < 0) throw Error("nimg must be >= 0");
        CnnModel &c = need_cnn(h);
        const size_t bytes = sizeof(float) * 4096 * (size_t)nimg;
        DevTmp dx(bytes), dy(bytes);
        HIPCHECK(hipMemcpy(dx.p, x, bytes, hipMemcpyHostToDevice));
        c.infer((const float *)dx.p, (float *)dy.p, nimg, h->stream);
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipMemcpy(out, dy.p, bytes, hipMemcpyDeviceToHost));
    });
}

int falcon_cnn_gradient(falcon_handle_t h, int nimg, const float *x, const float *gout, double *grad) {
    return guard([&] {
        if (nimg < 1) throw Error("nimg must be >= 1");
        CnnModel &c = need_cnn(h);
        const size_t bytes = sizeof(float) * 4096 * (size_t)nimg;
        DevTmp dx(bytes), dg(bytes);
        HIPCHECK(hipMemcpy(dx.p, x, bytes, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dg.p, gout, bytes, hipMemcpyHostToDevice));
        c.gradient((const float *)dx.p, (const float *)dg.p, nimg, grad, h->stream);
    });
}

int falcon_cnn_train(falcon_handle_t h, int64_t nimg, const float *x, const float *y, int epochs, int batch, double lr,
                     uint64_t seed, double *loss) {
    return guard([&] {
        if (nimg < 1 || epochs < 0 || batch < 1) throw Error("bad training sizes");
        CnnModel &c = need_cnn(h);
        const size_t bytes = sizeof(float) * 4096 * (size_t)nimg;
        DevTmp dx(bytes), dy(bytes);
        HIPCHECK(hipMemcpy(dx.p, x, bytes, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dy.p, y, bytes, hipMemcpyHostToDevice));
        const double l = c.train((const float *)dx.p, (const float *)dy.p, nimg, epochs, batch, lr, seed, h->stream);
        if (loss) *loss = l;
    });
}

int falcon_heat_train_corrector(falcon_handle_t h, int npairs, int blocks_per_pair, double dt_lo, double dt_hi,
                                int max_solves, int epochs, int batch, double lr, uint64_t seed, double *loss,
                                int64_t *pairs_made, double *seconds) {
    return guard([&] {
        HeatSolver &H = need_heat(h);
        if (npairs < 1 || blocks_per_pair < 1 || !(dt_lo > 0) || !(dt_hi >= dt_lo) || max_solves < 1)
            throw Error("bad training settings");
        CnnModel &c = need_cnn(h);
        const auto t0 = std::chrono::steady_clock::now();
        const size_t cap = (size_t)npairs * (size_t)blocks_per_pair;
        DevTmp dX(sizeof(float) * 4096 * cap), dY(sizeof(float) * 4096 * cap), dW(sizeof(double) * 4096 * cap),
            dS(sizeof(double) * cap), dRs(sizeof(int) * cap), dRb(sizeof(int) * cap);
        HeatRecord rec;
        rec.capacity = (long long)cap;
        rec.per_snapshot = blocks_per_pair;
        rec.rng = seed ^ 0x5851F42D4C957F2DULL;
        rec.X = (float *)dX.p; rec.Y = (float *)dY.p; rec.W = (double *)dW.p; rec.scale = (double *)dS.p;
        rec.rs = (int *)dRs.p; rec.rb = (int *)dRb.p;
        std::mt19937_64 rng(seed);
        std::uniform_real_distribution<double> U(std::log(dt_lo), std::log(dt_hi));
        std::normal_distribution<double> N01(0.0, 1.0);
        const int ns = H.nsys();
        std::vector<double> dts(ns), T(H.cells() * ns);
        for (int k = 0; k < max_solves && rec.count < rec.capacity; k++) {
            for (int s = 0; s < ns; s++) dts[s] = H.explicit_dt() * std::exp(U(rng));
            for (auto &v : T) v = 1e-3 * N01(rng);
            H.set_dt(dts);
            H.set_rhs_from_T(T);
            H.solve(1e-10, 300, kSolveRecord, &rec);
        }
        const double l = c.train(rec.X, rec.Y, rec.count, epochs, batch, lr, seed + 1, h->stream);
        H.set_corrector(&c);
        if (loss) *loss = l;
        if (pairs_made) *pairs_made = rec.pairs;
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}

int falcon_heat_set_graph(falcon_handle_t h, int on) {
    return guard([&] { need_heat(h, true).set_graph(on != 0); });
}

int falcon_heat_set_learned_blockstep(falcon_handle_t h, int on) {
    return guard([&] { need_heat(h).set_learned_blockstep(on != 0); });
}

int falcon_cnn_set_loss(falcon_handle_t h, double h1_alpha, double w_r, double w_t, double w_p, int scale_free) {
    return guard([&] {
        if (h1_alpha >= 0.0 && !(w_r >= 0.0 && w_t >= 0.0 && w_p >= 0.0))
            throw Error("loss face weights must be >= 0");
        if (h1_alpha >= 0.0 && scale_free && !(h1_alpha > 0.0))
            throw Error("the scale-free loss needs a positive mass weight");
        need_cnn(h).set_loss_h1(h1_alpha, w_r, w_t, w_p, scale_free != 0);
    });
}

int falcon_heat_set_learned_iters(falcon_handle_t h, int n) {
    return guard([&] { need_heat(h).set_learned_iters(n); });
}

int falcon_heat_set_coarse(falcon_handle_t h, double tol, int max_iter) {
    return guard([&] { need_heat(h, true).set_coarse(tol, max_iter); });
}

int falcon_heat_set_learned_sweeps(falcon_handle_t h, int pre0, int post0) {
    return guard([&] { need_heat(h).set_learned_sweeps(pre0, post0); });
}

int falcon_heat_attach_corrector(falcon_handle_t h) {
    return guard([&] { need_heat(h).set_corrector(&need_cnn(h)); });
}

int falcon_synchronize(falcon_handle_t h) {
    return guard([&] {
        if (h->stream) HIPCHECK(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
