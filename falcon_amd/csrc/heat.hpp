// Heat-conduction solver (heat.hip): multigrid-preconditioned flexible CG.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "layout.hpp"

namespace falcon {

struct HLevView {
    int m, nb;
    long long P, M3;
    int nsys;
};

struct HeatResult {
    std::vector<int> iters;
    std::vector<double> relres;
    double seconds = 0.0;
};

// Optional correction applied after the V-cycle inside the preconditioner:
// given the residual left by the V-cycle (interior, system-major) it adds its
// estimate of the remaining error to z (padded, system-major), on `stream`.
class HeatCorrector {
  public:
    virtual ~HeatCorrector() {}
    virtual void correct(const double *resid, double *z, hipStream_t stream) = 0;
};

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
    void apply(const std::vector<double> &x, std::vector<double> &y);
    // z = M r, one V-cycle, system-major interior vectors (parity tests)
    void vcycle_host(const std::vector<double> &r, std::vector<double> &z);
    HeatResult solve(double tol, int max_iter, bool use_corrector);
    void get_solution(int s, std::vector<double> &T) const;
    void get_rhs(int s, std::vector<double> &b) const;
    void set_corrector(HeatCorrector *c);
    // residual left by the V-cycle for a given right-hand side (training pairs)
    void vcycle_residual(const double *r_dev, double *z_dev_padded, double *resid_dev);
    double *device_solution();
    double *device_rhs();
    const HLevView level0_view() const;

  private:
    struct Impl;
    Impl *d;
};

}  // namespace falcon
