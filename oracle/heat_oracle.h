/* Heat-conduction oracle declarations (implemented in falcon_oracle.c).
 * TEST INFRASTRUCTURE ONLY. */
#ifndef FALCON_HEAT_ORACLE_H
#define FALCON_HEAT_ORACLE_H
typedef struct orc_heat orc_heat;
struct orc_ctx;
orc_heat *orc_heat_create(struct orc_ctx *c, double kappa0, double r_ref_cgs);
void orc_heat_destroy(orc_heat *H);
double orc_heat_dt_explicit(const orc_heat *H);
int orc_heat_levels(const orc_heat *H);
/* y = A x, interior arrays nblocks x (m,m,m) Fortran order */
void orc_heat_apply(orc_heat *H, double dt, const double *x, double *y);
/* z = M r: one V-cycle, zero initial guess; interior arrays, Fortran order per block */
void orc_heat_vcycle(orc_heat *H, double dt, const double *r, double *z);
/* Jacobi diagonal V/dt + sum of face coefficients, interior Fortran order */
void orc_heat_diag(orc_heat *H, double dt, double *d);
/* coarsest-level CG of the V-cycle: relative stop tol on r.z^(1/2), iteration cap */
void orc_heat_set_coarse(orc_heat *H, double tol, int max_iter);
/* zf += P zc, the V-cycle's prolongation from level l+1 to level l (test hook) */
void orc_heat_prolong(orc_heat *H, int l, const double *zc, double *zf);
int orc_heat_solve(orc_heat *H, double dt, const double *Tn, double tol, int max_iter, double *x_out, double *relres);
/* kpow: kappa = kappa0 (r/r_ref)^kpow, -3 or 0; top_dirichlet: value 0 on the top face */
orc_heat *orc_heat_create_opt(struct orc_ctx *c, double kappa0, double r_ref_cgs, int kpow, int top_dirichlet);
/* A x = b from x = 0 (dt = +inf: no V/dt term) */
int orc_heat_solve_b(orc_heat *H, double dt, const double *b, double tol, int max_iter, double *x_out, double *relres);
/* B = -grad(phi) (second order) with the PFSS boundary ghosts; see falcon_oracle.c */
void orc_pfss_field(orc_heat *H, const double *phi, const double *br, double *B);
#endif
