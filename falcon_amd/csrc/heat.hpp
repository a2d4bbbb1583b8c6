// Heat-conduction solver (heat.hip): multigrid-preconditioned flexible CG.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "layout.hpp"

namespace falcon {

struct HeatResult {
    std::vector<int> iters;
    std::vector<double> relres;
    double seconds = 0.0;
};

// Learned correction of the V-cycle: maps block images of the normalised
// residual left by the V-cycle to images of the normalised error estimate,
// x[img][4096] -> out[img][4096] with img = system * nblocks + block and the
// block's cells in [i][j][k] order (device pointers, on `stream`).
class HeatCorrector {
  public:
    virtual ~HeatCorrector() {}
    virtual void infer(const float *x, float *out, long long nimg, hipStream_t stream) = 0;
    // Solver form: image img = system * nb + block; the network sees
    // x / scale[system * scale_stride] and the result, times the same scale, is
    // written as FP64 to c[system * istride + block * 4096 + cell].  A zero
    // scale gives a zero image; systems with active[system] = 0 are skipped.
    virtual void infer_heat(const float *x, const double *scale, int scale_stride, int nb, long long istride, double *c,
                            long long nimg, const int *active, hipStream_t stream) = 0;
    // bring derived device data (packed weights) up to date before a captured solve
    virtual void prepare(hipStream_t stream) = 0;
};

// Training records collected during a solve (device arrays, capacity images):
// X network input images, W = x + z of the recorded V-cycle, Y targets (error
// left by the V-cycle over the feature scale), per record system, block, scale.
struct HeatRecord {
    long long capacity = 0, count = 0;   // block images
    long long pairs = 0;                 // snapshots: one system at one V-cycle application
    int per_snapshot = 16;               // blocks recorded per snapshot
    uint64_t rng = 0;
    float *X = nullptr, *Y = nullptr;
    double *W = nullptr, *scale = nullptr;
    int *rs = nullptr, *rb = nullptr;
};

enum { kSolvePlain = 0, kSolveLearned = 1, kSolveRecord = 2 };

class HeatSolver {
  public:
    HeatSolver();
    ~HeatSolver();
    void setup(const Layout &lay, const HeatParams &hp, int nsys, int nthreads);
    int levels() const;
    int nsys() const;
    int m() const;
    long long cells() const;                 // interior cells of one system
    hipStream_t stream() const;
    double explicit_dt() const;              // 1 / max(2 kappa sum_d 1/h_d^2)
    void set_dt(const std::vector<double> &dt);
    void set_rhs_from_T(const std::vector<double> &T);   // system-major, per block [i][j][k]
    void set_rhs(const std::vector<double> &b);          // the right-hand side itself, same layout
    // B = -grad(x) of system 0 at every finest-level cell, [block][3][i][j][k]
    // (PFSS field; bot: per block and bottom-face cell (j, k) the ghost offset
    // Br dr of the domain's bottom, geo: see build_pfss_tables)
    void pfss_field(const std::vector<double> &bot, const std::vector<double> &geo, std::vector<double> &B);
    void apply(const std::vector<double> &x, std::vector<double> &y);
    // z = M r, one V-cycle, system-major interior vectors (parity tests)
    void vcycle_host(const std::vector<double> &r, std::vector<double> &z);
    HeatResult solve(double tol, int max_iter, bool use_corrector);
    // mode kSolvePlain / kSolveLearned (needs a corrector) / kSolveRecord (plain
    // FCG that appends training records to rec)
    HeatResult solve(double tol, int max_iter, int mode, HeatRecord *rec);
    void get_solution(int s, std::vector<double> &T) const;
    void get_rhs(int s, std::vector<double> &b) const;
    void set_corrector(HeatCorrector *c);
    // finest-level smoothing sweeps of the V-cycle used with the learned
    // correction (recording and learned solves); the plain method keeps 2 / 2
    void set_learned_sweeps(int pre0, int post0);
    // coarsest-level CG of the V-cycle: stop at r.z <= tol^2 (r.z)_0 or after
    // max_iter iterations (defaults 1e-2, 50)
    void set_coarse(double tol, int max_iter);
    // plain and learned solves replay one captured FCG iteration (default) or
    // launch the same kernels eagerly
    void set_graph(bool on);
    // learned correction step per 16^3 block (a_b, default) or one per system
    void set_learned_blockstep(bool on);
    // the learned correction builds the first n directions of a solve, the
    // plain V-cycle the rest (0: every direction)
    void set_learned_iters(int n);

  private:
    struct Impl;
    Impl *d;
};

}  // namespace falcon
