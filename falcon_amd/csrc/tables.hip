// Ghost-table construction on the device: the donor search of every ghost
// cell (YinYangTree_FindNode, 3_6/ModYinYangTree.f90:449-505), the grouping
// of each block's ghosts into donor targets in first-appearance order
// (ModGC_GetGC_Targets_single, 3_6/ModGC.f90:53-147) and the trilinear base,
// weights and Yin-Yang rotation of every entry (3_6/ModMath.f90:265-280,
// 3_6/ModYinYang.f90:48-61).  The host supplies the per-block geometry rows and
// the transcendental rows (conv_rows: CoordConv and VecConv per (theta, phi)
// centre); every device step uses +, -, *, /, comparisons and floor only, in
// the host builder's order, so the tables equal Layout::build's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "layout.hpp"

namespace falcon {

namespace {

#define TCHECK(x)                                                                                           \
    do {                                                                                                    \
        hipError_t e_ = (x);                                                                                \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("device tables: ") + hipGetErrorString(e_)); \
    } while (0)

struct DNode {
    double rng[3][2];
    int yin, leaf, leaf_id, first_child;
    int N[3];
    int pad;
};

// the stack of the depth-first search: deeper trees are refused on the host
constexpr int kMaxDepth = 32;
// distinct donors of one block's ghosts (a 16^3 block has about 30)
constexpr int kMaxTargets = 256;

struct Geo {
    const double *xi, *xj, *xk;   // [block][NI], [block][NJ], [block][NK]
    const int *yin;               // [block]
    const double *conv;           // [block][NJ][NK][4]
    int ni, nj, nk, ng, NI, NJ, NK, nb;
};

__device__ __forceinline__ bool d_inside(const DNode &n, const double *rtp, const double *oth, int yin) {
    const double *h = (yin == n.yin) ? rtp : oth;
    return (h[0] - n.rng[0][0]) * (h[0] - n.rng[0][1]) <= 0.0 && (h[1] - n.rng[1][0]) * (h[1] - n.rng[1][1]) <= 0.0 &&
           (h[2] - n.rng[2][0]) * (h[2] - n.rng[2][1]) <= 0.0;
}

// Child number `it` of node n in Tree::find_branch's visiting order, or -1
// past the last.  Up to 8 children: in child order.  More: the children whose
// interval contains the point along each axis (at most 64 per axis, in axis
// order), visited as the product of the three lists.
__device__ int d_child(const DNode *nodes, const DNode &n, const double *h, int it) {
    const int N0 = n.N[0], N1 = n.N[1], N2 = n.N[2], nc = N0 * N1 * N2;
    if (nc <= 8) return it < nc ? n.first_child + it : -1;
    const int Ns[3] = {N0, N1, N2};
    int cnt[3];
    for (int a = 0; a < 3; a++) {
        int c = 0;
        for (int q = 0; q < Ns[a] && c < 64; q++) {
            const int ci = a == 0 ? q * N1 * N2 : (a == 1 ? q * N2 : q);
            const DNode &ch = nodes[n.first_child + ci];
            if ((h[a] - ch.rng[a][0]) * (h[a] - ch.rng[a][1]) <= 0.0) c++;
        }
        cnt[a] = c;
    }
    if (it >= cnt[0] * cnt[1] * cnt[2]) return -1;
    const int want[3] = {it / (cnt[1] * cnt[2]), (it / cnt[2]) % cnt[1], it % cnt[2]};
    int pick[3];
    for (int a = 0; a < 3; a++) {
        int c = 0;
        pick[a] = 0;
        for (int q = 0; q < Ns[a] && c < 64; q++) {
            const int ci = a == 0 ? q * N1 * N2 : (a == 1 ? q * N2 : q);
            const DNode &ch = nodes[n.first_child + ci];
            if ((h[a] - ch.rng[a][0]) * (h[a] - ch.rng[a][1]) <= 0.0) {
                if (c == want[a]) { pick[a] = q; break; }
                c++;
            }
        }
    }
    return n.first_child + (pick[0] * N1 + pick[1]) * N2 + pick[2];
}

// Tree::find_branch without recursion: the first child, in visiting order,
// that contains the point and whose subtree resolves it; -777 if none.
__device__ int d_find_branch(const DNode *nodes, int root, const double *rtp, const double *oth, int yin, int *err) {
    if (!d_inside(nodes[root], rtp, oth, yin)) return -777;
    int stk[kMaxDepth], itr[kMaxDepth];
    int sp = 1;
    stk[0] = root;
    itr[0] = 0;
    while (sp > 0) {
        const DNode &n = nodes[stk[sp - 1]];
        if (n.leaf) return n.leaf_id;
        const double *h = (yin == n.yin) ? rtp : oth;   // children share the parent's panel
        const int ch = d_child(nodes, n, h, itr[sp - 1]);
        if (ch < 0) { sp--; continue; }
        itr[sp - 1]++;
        if (d_inside(nodes[ch], rtp, oth, yin)) {
            if (sp == kMaxDepth) { *err = 1; return -777; }
            stk[sp] = ch;
            itr[sp] = 0;
            sp++;
        }
    }
    return -777;
}

// Donor of every ghost cell (ModGC_FindGC_Single, 3_6/ModGC.f90:37-50), one
// lane per padded cell: leaf id or -777; interior cells get 0.
__global__ void __launch_bounds__(256) k_find_donors(const DNode *__restrict__ nodes, Geo g,
                                                     int *__restrict__ donor, int *__restrict__ err) {
    const long long P = (long long)g.NI * g.NJ * g.NK;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P * g.nb) return;
    const int b = (int)(c / P);
    const long long q = c - (long long)b * P;
    const int I = (int)(q % g.NI), J = (int)((q / g.NI) % g.NJ), K = (int)(q / ((long long)g.NI * g.NJ));
    const int i = I - g.ng + 1, j = J - g.ng + 1, k = K - g.ng + 1;
    if (!(i < 1 || j < 1 || k < 1 || i > g.ni || j > g.nj || k > g.nk)) { donor[c] = 0; return; }
    const double rtp[3] = {g.xi[(long long)b * g.NI + I], g.xj[(long long)b * g.NJ + J], g.xk[(long long)b * g.NK + K]};
    const double *cv = g.conv + (((long long)b * g.NJ + J) * g.NK + K) * 4;
    const double oth[3] = {rtp[0], cv[0], cv[1]};
    const int yin = g.yin[b];
    // YinYangTree_FindNode: the Yin root first
    const int r = d_inside(nodes[0], rtp, oth, yin) ? d_find_branch(nodes, 0, rtp, oth, yin, err)
                                                    : d_find_branch(nodes, 1, rtp, oth, yin, err);
    donor[c] = r;
}

// Targets of each block (one lane per block): its ghosts in the scan order of
// ModGC_GetGC_Targets_single (i outer, k inner), grouped by donor in order of
// first appearance.  Out: ntg, the donors and counts, and the padded ordinals
// q (i + NI (j + NJ k)) of the ghosts in target order.
__global__ void __launch_bounds__(64) k_group_targets(Geo g, const int *__restrict__ donor, int *__restrict__ ntg,
                                                      int *__restrict__ tg_donor, int *__restrict__ tg_count,
                                                      int *__restrict__ qlist, int maxg, int *__restrict__ err) {
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= g.nb) return;
    const long long P = (long long)g.NI * g.NJ * g.NK;
    const int *D = donor + (long long)b * P;
    int *td = tg_donor + (long long)b * kMaxTargets, *tc = tg_count + (long long)b * kMaxTargets;
    int n = 0;
    for (int pass = 0; pass < 2; pass++) {
        int total = 0;
        if (pass == 1) {
            // offsets of the target lists (counts become running positions)
            for (int t = 0; t < n; t++) { const int c = tc[t]; tc[t] = total; total += c; }
        }
        for (int I = 0; I < g.NI; I++)
            for (int J = 0; J < g.NJ; J++)
                for (int K = 0; K < g.NK; K++) {
                    const int i = I - g.ng + 1, j = J - g.ng + 1, k = K - g.ng + 1;
                    if (!(i < 1 || i > g.ni || j < 1 || j > g.nj || k < 1 || k > g.nk)) continue;
                    const long long q = I + (long long)g.NI * (J + (long long)g.NJ * K);
                    const int d = D[q];
                    if (d == -777) continue;
                    int s = 0;
                    while (s < n && td[s] != d) s++;
                    if (pass == 0) {
                        if (s == n) {
                            if (n == kMaxTargets) { *err = 2; return; }
                            td[n] = d;
                            tc[n] = 0;
                            n++;
                        }
                        tc[s]++;
                    } else {
                        const int pos = tc[s]++;
                        if (pos >= maxg) { *err = 3; return; }
                        qlist[(long long)b * maxg + pos] = (int)q;
                    }
                }
    }
    // tc[t] now holds the end of list t: back to counts
    for (int t = n - 1; t > 0; t--) tc[t] -= tc[t - 1];
    ntg[b] = n;
}

// One lane per ghost entry in table order (blocks, then targets, then scan
// order): base cell, weights and rotation as Layout::build forms them.
__global__ void __launch_bounds__(256) k_entry_weights(Geo g, const long long *__restrict__ eoff,
                                                       const int *__restrict__ ntg, const int *__restrict__ tg_donor,
                                                       const int *__restrict__ tg_count, const int *__restrict__ qlist,
                                                       int maxg, GhostEntry *__restrict__ out) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= eoff[g.nb]) return;
    int lo = 0, hi = g.nb - 1;   // block: the last b with eoff[b] <= e
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (eoff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    const int b = lo;
    const int r = (int)(e - eoff[b]);
    const int *td = tg_donor + (long long)b * kMaxTargets, *tc = tg_count + (long long)b * kMaxTargets;
    int t = 0, acc = 0;
    while (t < ntg[b] - 1 && acc + tc[t] <= r) { acc += tc[t]; t++; }
    const int d = td[t] - 1;
    const int q = qlist[(long long)b * maxg + r];
    const int I = q % g.NI, J = (q / g.NI) % g.NJ, K = q / (g.NI * g.NJ);
    const int ng = g.ng;
    double x[3] = {g.xi[(long long)b * g.NI + I], g.xj[(long long)b * g.NJ + J], g.xk[(long long)b * g.NK + K]};
    const bool yy = g.yin[d] != g.yin[b];
    const double *cv = g.conv + (((long long)b * g.NJ + J) * g.NK + K) * 4;
    if (yy) { x[1] = cv[0]; x[2] = cv[1]; }
    const double *sxi = g.xi + (long long)d * g.NI, *sxj = g.xj + (long long)d * g.NJ, *sxk = g.xk + (long long)d * g.NK;
    const double dxi = sxi[ng + 1] - sxi[ng], dxj = sxj[ng + 1] - sxj[ng], dxk = sxk[ng + 1] - sxk[ng];
    const double base = (double)((float)(-ng) + 1.0f);
    const double lim[3] = {(double)((float)g.ni + 0.0f), (double)((float)g.nj + 0.0f), (double)((float)g.nk + 0.0f)};
    const double p[3] = {base + (x[0] - sxi[0]) / dxi, base + (x[1] - sxj[0]) / dxj, base + (x[2] - sxk[0]) / dxk};
    GhostEntry o;
    o.recv_block = b;
    o.gi = I - ng + 1; o.gj = J - ng + 1; o.gk = K - ng + 1;
    o.donor = d;
    int bi[3];
    for (int a = 0; a < 3; a++) {
        double v = p[a] < lim[a] ? p[a] : lim[a];
        v = v > 1.0 ? v : 1.0;
        bi[a] = (int)floor(v);
        o.w[a] = v - (double)bi[a];
    }
    o.i0 = bi[0]; o.j0 = bi[1]; o.k0 = bi[2];
    o.yy = yy ? 1 : 0;
    o.rot[0] = yy ? cv[2] : 0.0;
    o.rot[1] = yy ? cv[3] : 0.0;
    out[e] = o;
}

struct DBuf {
    void *p = nullptr;
    ~DBuf() { if (p) (void)hipFree(p); }
    template <typename T>
    T *alloc(size_t n) {
        TCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        return (T *)p;
    }
};

}  // namespace

void build_ghosts_device(Layout &lay, int nthreads, hipStream_t s) {
    Tree &t = lay.tree;
    lay.build_geom();
    const int nb = t.nleaf, ni = t.ni, nj = t.nj, nk = t.nk, ng = t.ng;
    const int NI = ni + 2 * ng, NJ = nj + 2 * ng, NK = nk + 2 * ng;
    const long long P = (long long)NI * NJ * NK;
    const int maxg = (int)(P - (long long)ni * nj * nk);

    // nodes, depth-checked
    std::vector<DNode> nodes(t.nodes.size());
    std::vector<int> depth(t.nodes.size(), 0);
    for (size_t n = 0; n < t.nodes.size(); n++) {
        const Node &x = t.nodes[n];
        DNode &y = nodes[n];
        for (int a = 0; a < 3; a++) { y.rng[a][0] = x.rng[a][0]; y.rng[a][1] = x.rng[a][1]; y.N[a] = x.N[a]; }
        y.yin = x.yin; y.leaf = x.leaf; y.leaf_id = x.leaf_id; y.first_child = x.first_child; y.pad = 0;
    }
    for (size_t n = 0; n < t.nodes.size(); n++) {   // children are stored after their parent
        const Node &x = t.nodes[n];
        if (x.leaf) continue;
        const int nc = x.N[0] * x.N[1] * x.N[2];
        for (int c = 0; c < nc; c++) depth[x.first_child + c] = depth[n] + 1;
    }
    if (*std::max_element(depth.begin(), depth.end()) >= kMaxDepth) throw std::runtime_error("device tables: tree too deep");

    // geometry and transcendental rows (host)
    std::vector<double> xi((size_t)nb * NI), xj((size_t)nb * NJ), xk((size_t)nb * NK), conv;
    std::vector<int> yin(nb);
    for (int b = 0; b < nb; b++) {
        const BlockGeom &gg = lay.geom[b];
        std::copy(gg.xi.begin(), gg.xi.end(), xi.begin() + (size_t)b * NI);
        std::copy(gg.xj.begin(), gg.xj.end(), xj.begin() + (size_t)b * NJ);
        std::copy(gg.xk.begin(), gg.xk.end(), xk.begin() + (size_t)b * NK);
        yin[b] = gg.yin;
    }
    conv_rows(lay, nthreads, conv);

    DBuf b_nodes, b_xi, b_xj, b_xk, b_yin, b_conv, b_donor, b_err, b_ntg, b_tgd, b_tgc, b_q, b_eoff, b_out;
    DNode *d_nodes = b_nodes.alloc<DNode>(nodes.size());
    Geo g{b_xi.alloc<double>(xi.size()), b_xj.alloc<double>(xj.size()), b_xk.alloc<double>(xk.size()),
          b_yin.alloc<int>(nb), b_conv.alloc<double>(conv.size()), ni, nj, nk, ng, NI, NJ, NK, nb};
    TCHECK(hipMemcpyAsync(d_nodes, nodes.data(), sizeof(DNode) * nodes.size(), hipMemcpyHostToDevice, s));
    TCHECK(hipMemcpyAsync((void *)g.xi, xi.data(), sizeof(double) * xi.size(), hipMemcpyHostToDevice, s));
    TCHECK(hipMemcpyAsync((void *)g.xj, xj.data(), sizeof(double) * xj.size(), hipMemcpyHostToDevice, s));
    TCHECK(hipMemcpyAsync((void *)g.xk, xk.data(), sizeof(double) * xk.size(), hipMemcpyHostToDevice, s));
    TCHECK(hipMemcpyAsync((void *)g.yin, yin.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
    TCHECK(hipMemcpyAsync((void *)g.conv, conv.data(), sizeof(double) * conv.size(), hipMemcpyHostToDevice, s));
    int *d_donor = b_donor.alloc<int>((size_t)nb * P);
    int *d_err = b_err.alloc<int>(1);
    TCHECK(hipMemsetAsync(d_err, 0, sizeof(int), s));
    const long long ncell = (long long)nb * P;
    hipLaunchKernelGGL(k_find_donors, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, d_nodes, g, d_donor, d_err);
    TCHECK(hipGetLastError());
    int *d_ntg = b_ntg.alloc<int>(nb);
    int *d_tgd = b_tgd.alloc<int>((size_t)nb * kMaxTargets), *d_tgc = b_tgc.alloc<int>((size_t)nb * kMaxTargets);
    int *d_q = b_q.alloc<int>((size_t)nb * maxg);
    hipLaunchKernelGGL(k_group_targets, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, s, g, (const int *)d_donor, d_ntg,
                       d_tgd, d_tgc, d_q, maxg, d_err);
    TCHECK(hipGetLastError());
    int err = 0;
    std::vector<int> ntg(nb), tgd((size_t)nb * kMaxTargets), tgc((size_t)nb * kMaxTargets);
    TCHECK(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
    TCHECK(hipMemcpyAsync(ntg.data(), d_ntg, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    TCHECK(hipMemcpyAsync(tgd.data(), d_tgd, sizeof(int) * tgd.size(), hipMemcpyDeviceToHost, s));
    TCHECK(hipMemcpyAsync(tgc.data(), d_tgc, sizeof(int) * tgc.size(), hipMemcpyDeviceToHost, s));
    TCHECK(hipStreamSynchronize(s));
    if (err) throw std::runtime_error("device tables: search limits exceeded (code " + std::to_string(err) + ")");

    // entry offsets per block
    std::vector<long long> eoff(nb + 1, 0);
    for (int b = 0; b < nb; b++) {
        long long n = 0;
        for (int q = 0; q < ntg[b]; q++) n += tgc[(size_t)b * kMaxTargets + q];
        eoff[b + 1] = eoff[b] + n;
    }
    const long long ne = eoff[nb];
    long long *d_eoff = b_eoff.alloc<long long>(nb + 1);
    GhostEntry *d_out = b_out.alloc<GhostEntry>((size_t)ne);
    TCHECK(hipMemcpyAsync(d_eoff, eoff.data(), sizeof(long long) * (nb + 1), hipMemcpyHostToDevice, s));
    if (ne > 0)
        hipLaunchKernelGGL(k_entry_weights, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, g,
                           (const long long *)d_eoff, (const int *)d_ntg, (const int *)d_tgd, (const int *)d_tgc,
                           (const int *)d_q, maxg, d_out);
    TCHECK(hipGetLastError());

    // results into the layout
    std::vector<int> q((size_t)nb * maxg);
    lay.ghosts.resize((size_t)ne);
    TCHECK(hipMemcpyAsync(q.data(), d_q, sizeof(int) * q.size(), hipMemcpyDeviceToHost, s));
    if (ne > 0)
        TCHECK(hipMemcpyAsync(lay.ghosts.data(), d_out, sizeof(GhostEntry) * (size_t)ne, hipMemcpyDeviceToHost, s));
    std::vector<int> donor;
    if (lay.keep_gcib) {
        donor.resize((size_t)nb * P);
        TCHECK(hipMemcpyAsync(donor.data(), d_donor, sizeof(int) * donor.size(), hipMemcpyDeviceToHost, s));
    }
    TCHECK(hipStreamSynchronize(s));
    lay.targets.assign(nb, {});
    for (int b = 0; b < nb; b++) {
        auto &tg = lay.targets[b];
        int pos = 0;
        for (int t2 = 0; t2 < ntg[b]; t2++) {
            const int n = tgc[(size_t)b * kMaxTargets + t2];
            tg.push_back({tgd[(size_t)b * kMaxTargets + t2],
                          std::vector<int32_t>(q.begin() + (size_t)b * maxg + pos, q.begin() + (size_t)b * maxg + pos + n)});
            pos += n;
        }
    }
    if (lay.keep_gcib) {
        lay.gcib.assign(nb, {});
        for (int b = 0; b < nb; b++) {
            std::vector<int32_t> v(donor.begin() + (size_t)b * P, donor.begin() + (size_t)(b + 1) * P);
            // interior cells: -777 as the host builder leaves them
            for (int K = 0; K < NK; K++)
                for (int J = 0; J < NJ; J++)
                    for (int I = 0; I < NI; I++) {
                        const int i = I - ng + 1, j = J - ng + 1, k = K - ng + 1;
                        if (!(i < 1 || j < 1 || k < 1 || i > ni || j > nj || k > nk))
                            v[I + (size_t)NI * (J + (size_t)NJ * K)] = -777;
                    }
            lay.gcib[b] = std::move(v);
        }
    }
}

}  // namespace falcon
