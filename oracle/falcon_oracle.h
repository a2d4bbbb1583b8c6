/* Oracle: CPU FP64 restatement of the Falcon reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or the timed
 * CPU baseline; the product (falcon_amd/) never links or calls it.
 *
 * Arrays follow the reference's Fortran layout: padded block arrays are
 * (-ng+1:ni+ng, -ng+1:nj+ng, -ng+1:nk+ng, nvar) column-major with r fastest,
 * interior arrays (1:ni,1:nj,1:nk,nvar).  Variables in 3_6 order without psi:
 * rho1, vr, vt, vp, br, bt, bp, s1 (3_6/ModBlock.f90:153-163).
 */
#ifndef FALCON_ORACLE_H
#define FALCON_ORACLE_H
#include <stdint.h>

#define ORC_NVAR 8

typedef struct orc_ctx orc_ctx;

/* ---- tree / layout (3_6/ModYinYangTree.f90, 3_6/ModAMR.f90) */
orc_ctx *orc_create(int ni, int nj, int nk, int ng, double r0_rsun, double r1_rsun);
void orc_destroy(orc_ctx *c);
/* Divide every leaf into nr x nt x np children with the split rule of
 * YinYangTree_Divide (3_6/ModYinYangTree.f90:239-255); N=1 or 2 per axis is the
 * reference's own case, other N generalise the same formulas. */
void orc_divide_all(orc_ctx *c, int nr, int nt, int np);
/* One ModAMR_Divide_OneBranch pass (3_6/ModAMR.f90:65-99) on both panels. */
void orc_amr(orc_ctx *c, const int flags[3], int if_global, const double rng_rsun[6]);
/* split the listed leaves (0-based depth-first ordinals, Yin then Yang) 2x2x2; -1 on a bad ordinal */
int orc_refine(orc_ctx *c, int n, const int *ordinals);
/* Label leaves, build blocks, geometry and ghost tables. Returns leaf count. */
int orc_finalize(orc_ctx *c);

int orc_nblocks(const orc_ctx *c);
int orc_nleaf_yin(const orc_ctx *c);
/* range[6] = (r lo, t lo, p lo, r hi, t hi, p hi) as xijk_range(3,2); flags[3]=yin,top,bottom */
void orc_block_info(const orc_ctx *c, int ib, double range[6], int flags[3]);
void orc_block_centres(const orc_ctx *c, int ib, double *xi, double *xj, double *xk);
/* GC_iBlocks_III of one block, padded, Fortran order; -777 outside domain, -1 interior */
void orc_gc_iblocks(const orc_ctx *c, int ib, int32_t *out);
int orc_ntargets(const orc_ctx *c, int ib);
/* target it of block ib: donor leaf id (1-based) and nGC; ijk (nGC x 3, column-major) and xijk */
int orc_target(const orc_ctx *c, int ib, int it, int32_t *ijk, double *xijk);
int orc_target_donor(const orc_ctx *c, int ib, int it);
int orc_nsources(const orc_ctx *c, int ib);
int orc_source(const orc_ctx *c, int ib, int is, int32_t *recv_block, int32_t *ijk, double *xijk);

/* ---- data */
/* bg: (ni+2ng) x 8 column-major: g, rho0, p0, T0, gamma1, gamma3, heat, xi_rsst */
void orc_set_background(orc_ctx *c, int ib, const double *bg);
void orc_set_state(orc_ctx *c, int ib, const double *u_padded);
void orc_get_state(const orc_ctx *c, int ib, double *u_padded);

/* ---- operators */
void orc_exchange(orc_ctx *c, int which);          /* 0: U, 1: U_rk  (SendRecvGC_new) */
void orc_rhs(orc_ctx *c, int ib, double *r_out);    /* BC + ModEquation_Dynamo on U */
void orc_diffusion(orc_ctx *c, int ib, double *r_out); /* ModDiffusion_Aritificial_1(h=2) on U */
double orc_dt(orc_ctx *c, double cfl);               /* ModWaveSpeed + ModTimeStep */
double orc_step(orc_ctx *c, double cfl);             /* one ModAdvance_rk4; returns dt */
/* Pin of orc_step's composition: on != 0 replaces the stage right-hand side by
 * lambda * (stage input) and the diffusion's by mu * U (interior, every
 * variable), leaving the stage buffers, coefficients, exchanges and dt as they
 * are; one step then multiplies each interior value by
 * (1 + z + z^2/2 + z^3/6 + z^4/24) (1 + mu dt/2), z = lambda dt. */
void orc_set_linear_rhs(orc_ctx *c, int on, double lambda, double mu);
/* GLM divergence cleaning (3_6/ModDivB.f90:13-38, DivB_option=1): on != 0
 * makes psi a ninth variable of the state, stepped, exchanged and diffused
 * with the others; psi is set and read as one padded array per block. */
void orc_set_glm(orc_ctx *c, int on);
void orc_set_psi(orc_ctx *c, int ib, const double *psi_padded);
void orc_get_psi(const orc_ctx *c, int ib, double *psi_padded);
void orc_set_threads(int n);
void orc_coordconv(const double in[3], double out[3]);
void orc_vecconv_coeffs(const double x[3], double out[2]);

/* ---- heat conduction oracle (heat_oracle.c) */
#include "heat_oracle.h"
/* ---- CNN oracle (cnn_oracle.c) */
#include "cnn_oracle.h"

#endif
