/* Oracle: CPU FP64 restatement of the Falcon reference RK4 hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see falcon_oracle.h).  Each function cites the
 * reference routine it restates; operation order and the single-precision
 * literals of the reference (3_6/ModConst.f90, `1./12.`, `dpi*0.25`, ...) are
 * reproduced so results can be compared with the flang build of the reference.
 * Compiled with -ffp-contract=off: no fused multiply-add anywhere.
 */
#define _GNU_SOURCE
#include "falcon_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 3_6/ModConst.f90: default-real (float) parameters */
static const float DPI_F = 3.1415926535f;
static const float RSUN_F = 6.955e10f;

typedef struct onode {
    int leaf, yin, ileaf, nch;
    double rng[3][2];
    struct onode *ch;
} onode;

typedef struct {
    int donor;          /* 1-based leaf id */
    int ngc;
    int32_t *ijk;       /* ngc x 3 column-major (Fortran ijk_list) */
    double *xijk;       /* ngc x 3 column-major, receiver frame */
} otarget;

typedef struct {
    int recv;           /* 1-based leaf id of the receiving block */
    int recv_yin;
    int ngc;
    int32_t *ijk;
    double *xijk;       /* donor frame */
} osource;

typedef struct {
    int iblock, yin, top, bottom;
    double rng[3][2];
    double dxi, dxj, dxk;
    double *xi, *xj, *xk;   /* padded centres, index 0 <-> -ng+1 */
    double *h;              /* interior h_LLL */
    int32_t *gcib;          /* padded GC_iBlocks_III */
    int ntg; otarget *tg;
    int nsc; osource *sc;
    /* background 1-D (padded in r) */
    double *gamma1, *gm1, *rho0, *g_over_rho0, *p0_over_rho0, *rho0T0, *heat, *xirsst;
    double *u, *urk;        /* padded x 9 (the ninth, psi, used with GLM only) */
    double *r;              /* interior x 9 */
    double *vw;             /* interior v_wave of the step (GLM only) */
} oblock;

struct orc_ctx {
    int ni, nj, nk, ng;
    int NI, NJ, NK;          /* padded extents */
    double r_range[2];
    onode yin, yang;
    int nleaf, nleaf_yin;
    oblock *b;
    /* pin of the step composition (orc_set_linear_rhs): when lin_on, orc_step
     * takes R = lin_lambda * (stage input) for the stages and R = lin_mu * U
     * for the diffusion, in place of ModEquation_Dynamo / ModDiffusion */
    int lin_on;
    double lin_lambda, lin_mu;
    /* variables stepped: 8, or 9 with GLM divergence cleaning (psi last) */
    int nv;
};

#define PAD(c, i, j, k) ((size_t)((i) + (c)->ng - 1) + (size_t)(c)->NI * ((size_t)((j) + (c)->ng - 1) + (size_t)(c)->NJ * (size_t)((k) + (c)->ng - 1)))
#define INT(c, i, j, k) ((size_t)((i) - 1) + (size_t)(c)->ni * ((size_t)((j) - 1) + (size_t)(c)->nj * (size_t)((k) - 1)))

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* ------------------------------------------------------------------ tree */

static void node_init(onode *n, double rng[3][2], int yin) {
    memset(n, 0, sizeof(*n));
    n->leaf = 1; n->yin = yin;
    memcpy(n->rng, rng, sizeof(n->rng));
}

orc_ctx *orc_create(int ni, int nj, int nk, int ng, double r0_rsun, double r1_rsun) {
    orc_ctx *c = calloc(1, sizeof(orc_ctx));
    c->ni = ni; c->nj = nj; c->nk = nk; c->ng = ng;
    c->nv = 8;
    c->NI = ni + 2 * ng; c->NJ = nj + 2 * ng; c->NK = nk + 2 * ng;
    /* r_range = r_range_Rsun*R_sun__CGS (3_6/ModReadParameters.f90:71) */
    c->r_range[0] = r0_rsun * (double)RSUN_F;
    c->r_range[1] = r1_rsun * (double)RSUN_F;
    /* YinYangTree_InitTree (3_6/ModYinYangTree.f90:91-110): float products */
    double rng[3][2];
    rng[0][0] = c->r_range[0]; rng[0][1] = c->r_range[1];
    rng[1][0] = (double)(DPI_F * 0.25f); rng[1][1] = (double)(DPI_F * 0.75f);
    rng[2][0] = (double)(-DPI_F * 0.75f); rng[2][1] = (double)(DPI_F * 0.75f);
    node_init(&c->yin, rng, 1);
    node_init(&c->yang, rng, 0);
    return c;
}

static void node_free(onode *n) {
    if (!n->leaf) {
        for (int i = 0; i < n->nch; i++) node_free(&n->ch[i]);
        free(n->ch);
    }
}

/* YinYangTree_Divide (3_6/ModYinYangTree.f90:210-269) with N_rtp per axis.
 * Exponents and fractions are default-real expressions ([N-ir+1.0,...]/N). */
static void node_divide(onode *n, const int N[3]) {
    if (!n->leaf) return;
    n->nch = N[0] * N[1] * N[2];
    n->ch = calloc((size_t)n->nch, sizeof(onode));
    int ic = 0;
    for (int ir = 1; ir <= N[0]; ir++)
        for (int it = 1; it <= N[1]; it++)
            for (int ip = 1; ip <= N[2]; ip++) {
                double r[3][2];
                float ea0 = ((float)(N[0] - ir) + 1.0f) / (float)N[0];
                float ea1 = ((float)(N[0] - ir) + 0.0f) / (float)N[0];
                float eb0 = ((float)(-ir) + 1.0f) / (float)N[0];
                float eb1 = ((float)(-ir) + 0.0f) / (float)N[0];
                r[0][0] = pow(n->rng[0][0], (double)ea0) / pow(n->rng[0][1], (double)eb0);
                r[0][1] = pow(n->rng[0][0], (double)ea1) / pow(n->rng[0][1], (double)eb1);
                float ta0 = (float)(N[1] - it) + 1.0f, ta1 = (float)(N[1] - it) + 0.0f;
                float tb0 = (float)(-it) + 1.0f, tb1 = (float)(-it) + 0.0f;
                r[1][0] = (n->rng[1][0] * (double)ta0) / (double)N[1] - (n->rng[1][1] * (double)tb0) / (double)N[1];
                r[1][1] = (n->rng[1][0] * (double)ta1) / (double)N[1] - (n->rng[1][1] * (double)tb1) / (double)N[1];
                float pa0 = (float)(N[2] - ip) + 1.0f, pa1 = (float)(N[2] - ip) + 0.0f;
                float pb0 = (float)(-ip) + 1.0f, pb1 = (float)(-ip) + 0.0f;
                r[2][0] = (n->rng[2][0] * (double)pa0) / (double)N[2] - (n->rng[2][1] * (double)pb0) / (double)N[2];
                r[2][1] = (n->rng[2][0] * (double)pa1) / (double)N[2] - (n->rng[2][1] * (double)pb1) / (double)N[2];
                node_init(&n->ch[ic++], r, n->yin);
            }
    n->leaf = 0;
}

static void divide_branch(onode *n, const int N[3]) {
    if (n->leaf) node_divide(n, N);
    else for (int i = 0; i < n->nch; i++) divide_branch(&n->ch[i], N);
}

void orc_divide_all(orc_ctx *c, int nr, int nt, int np) {
    int N[3] = {nr, nt, np};
    divide_branch(&c->yin, N);
    divide_branch(&c->yang, N);
}

/* ModMath_IfLinesInterSect (3_6/ModMath.f90:5-16): strict inequalities */
static int lines_intersect(const double a[2], const double b[2]) {
    return (a[0] - b[0]) * (a[0] - b[1]) < 0.0 || (a[1] - b[0]) * (a[1] - b[1]) < 0.0 ||
           (b[0] - a[0]) * (b[0] - a[1]) < 0.0 || (b[1] - a[0]) * (b[1] - a[1]) < 0.0;
}

/* ModAMR_Divide_OneBranch (3_6/ModAMR.f90:65-99) */
static void amr_branch(onode *n, const int N[3], int global, const double rr[2], const double tr[2], const double pr[2]) {
    if (n->leaf) {
        if (global) {
            if (lines_intersect(n->rng[0], rr)) node_divide(n, N);
        } else if (n->yin) {
            if (lines_intersect(n->rng[0], rr) && lines_intersect(n->rng[1], tr) && lines_intersect(n->rng[2], pr))
                node_divide(n, N);
        }
    } else {
        for (int i = 0; i < n->nch; i++) amr_branch(&n->ch[i], N, global, rr, tr, pr);
    }
}

void orc_amr(orc_ctx *c, const int flags[3], int if_global, const double rng_rsun[6]) {
    int N[3] = {flags[0] ? 2 : 1, flags[1] ? 2 : 1, flags[2] ? 2 : 1};
    /* AMR1%rtp_range_Rsun(1,:)*R_sun__CGS; angles used as given */
    double rr[2] = {rng_rsun[0] * (double)RSUN_F, rng_rsun[3] * (double)RSUN_F};
    double tr[2] = {rng_rsun[1], rng_rsun[4]};
    double pr[2] = {rng_rsun[2], rng_rsun[5]};
    amr_branch(&c->yin, N, if_global, rr, tr, pr);
    amr_branch(&c->yang, N, if_global, rr, tr, pr);
}

/* local refinement: split the listed leaves (0-based depth-first ordinals, Yin then Yang) 2x2x2 */
static void leaves_dfs(onode *n, onode **out, int *k) {
    if (n->leaf) out[(*k)++] = n;
    else for (int i = 0; i < n->nch; i++) leaves_dfs(&n->ch[i], out, k);
}
static int count_leaves(const onode *n) {
    if (n->leaf) return 1;
    int s = 0;
    for (int i = 0; i < n->nch; i++) s += count_leaves(&n->ch[i]);
    return s;
}

int orc_refine(orc_ctx *c, int n, const int *ordinals) {
    int total = count_leaves(&c->yin) + count_leaves(&c->yang), k = 0;
    onode **lv = malloc(sizeof(onode *) * (size_t)total);
    leaves_dfs(&c->yin, lv, &k);
    leaves_dfs(&c->yang, lv, &k);
    const int N[3] = {2, 2, 2};
    for (int i = 0; i < n; i++) {
        if (ordinals[i] < 0 || ordinals[i] >= total) { free(lv); return -1; }
        node_divide(lv[ordinals[i]], N);
    }
    free(lv);
    return 0;
}

static void label(orc_ctx *c, onode *n) {
    if (n->leaf) n->ileaf = ++c->nleaf;
    else for (int i = 0; i < n->nch; i++) label(c, &n->ch[i]);
}

static void collect(orc_ctx *c, onode *n) {
    if (n->leaf) {
        oblock *b = &c->b[n->ileaf - 1];
        b->iblock = n->ileaf; b->yin = n->yin;
        memcpy(b->rng, n->rng, sizeof(b->rng));
    } else for (int i = 0; i < n->nch; i++) collect(c, &n->ch[i]);
}

/* ModYinYang_CoordConv_0D (3_6/ModYinYang.f90:8-17).  The reference takes sin
 * and cos of the same angle, which flang (like any LLVM compiler) evaluates as
 * one glibc sincos call; sincos can differ from cos by one ulp (theta =
 * 1.0983302628155798: cos 0x3FDD2016D70B5643, sincos 0x3FDD2016D70B5642), so
 * the pairs are taken from sincos here as well. */
static void coordconv(const double in[3], double out[3]) {
    double st, ct, sp, cp;
    sincos(in[1], &st, &ct);
    sincos(in[2], &sp, &cp);
    out[0] = in[0];
    out[1] = acos(st * sp);
    double a = acos(-((st * cp) / sqrt((st * st) * (cp * cp) + ct * ct)));
    out[2] = copysign(fabs(a), ct);
}

/* YinYangTree_IfInsideNode (3_6/ModYinYangTree.f90:424-445) */
static int inside(const onode *n, const double rtp[3], int yin) {
    double h[3];
    if (yin == n->yin) { h[0] = rtp[0]; h[1] = rtp[1]; h[2] = rtp[2]; }
    else coordconv(rtp, h);
    return (h[0] - n->rng[0][0]) * (h[0] - n->rng[0][1]) <= 0.0 &&
           (h[1] - n->rng[1][0]) * (h[1] - n->rng[1][1]) <= 0.0 &&
           (h[2] - n->rng[2][0]) * (h[2] - n->rng[2][1]) <= 0.0;
}

/* YinYangTree_FindNode_OneBranch (3_6/ModYinYangTree.f90:468-505) */
static const onode *find_branch(const onode *n, const double rtp[3], int yin) {
    if (!inside(n, rtp, yin)) return NULL;
    if (n->leaf) return n;
    for (int i = 0; i < n->nch; i++) {
        const onode *r = NULL;
        if (inside(&n->ch[i], rtp, yin)) r = find_branch(&n->ch[i], rtp, yin);
        if (r) return r;
    }
    return NULL;
}

/* YinYangTree_FindNode (3_6/ModYinYangTree.f90:449-464): Yin root tested first */
static const onode *find_node(const orc_ctx *c, const double rtp[3], int yin) {
    if (inside(&c->yin, rtp, yin)) return find_branch(&c->yin, rtp, yin);
    return find_branch(&c->yang, rtp, yin);
}

/* ModBlock_InitGrid (3_6/ModBlock.f90:288-438), spherical branch */
static void init_grid(orc_ctx *c, oblock *b) {
    int ni = c->ni, nj = c->nj, nk = c->nk, ng = c->ng;
    b->xi = malloc(sizeof(double) * c->NI);
    b->xj = malloc(sizeof(double) * c->NJ);
    b->xk = malloc(sizeof(double) * c->NK);
    for (int i = 1 - ng; i <= ni + ng; i++)
        b->xi[i + ng - 1] = (b->rng[0][0] * (double)((float)(ni - i) + 0.5f) + b->rng[0][1] * (double)((float)i - 0.5f)) / (double)ni;
    for (int j = 1 - ng; j <= nj + ng; j++)
        b->xj[j + ng - 1] = (b->rng[1][0] * (double)((float)(nj - j) + 0.5f) + b->rng[1][1] * (double)((float)j - 0.5f)) / (double)nj;
    for (int k = 1 - ng; k <= nk + ng; k++)
        b->xk[k + ng - 1] = (b->rng[2][0] * (double)((float)(nk - k) + 0.5f) + b->rng[2][1] * (double)((float)k - 0.5f)) / (double)nk;
    b->dxi = (b->rng[0][1] - b->rng[0][0]) / (double)ni;
    b->dxj = (b->rng[1][1] - b->rng[1][0]) / (double)nj;
    b->dxk = (b->rng[2][1] - b->rng[2][0]) / (double)nk;
    b->h = malloc(sizeof(double) * (size_t)ni * nj * nk);
    for (int k = 1; k <= nk; k++)
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double xi = b->xi[i + ng - 1];
                double a = b->dxi, bb = xi * b->dxj, cc = (xi * sin(b->xj[j + ng - 1])) * b->dxk;
                double m = a < bb ? a : bb;
                m = m < cc ? m : cc;
                b->h[INT(c, i, j, k)] = m;
            }
    /* YinYangTree_if_top_bottom (3_6/ModYinYangTree.f90:273-288) */
    b->top = b->xi[ni + 2 * ng - 1] > c->r_range[1];
    b->bottom = b->xi[0] < c->r_range[0];
}

/* ModGC_FindGC_Single + ModGC_GetGC_Targets_single (3_6/ModGC.f90:25-147) */
static void find_gc(orc_ctx *c, oblock *b) {
    int ni = c->ni, nj = c->nj, nk = c->nk, ng = c->ng;
    size_t np = (size_t)c->NI * c->NJ * c->NK;
    b->gcib = malloc(sizeof(int32_t) * np);
    for (size_t q = 0; q < np; q++) b->gcib[q] = -777;
    for (int i = 1 - ng; i <= ni + ng; i++)
        for (int j = 1 - ng; j <= nj + ng; j++)
            for (int k = 1 - ng; k <= nk + ng; k++) {
                if (i < 1 || j < 1 || k < 1 || i > ni || j > nj || k > nk) {
                    double rtp[3] = {b->xi[i + ng - 1], b->xj[j + ng - 1], b->xk[k + ng - 1]};
                    const onode *n = find_node(c, rtp, b->yin);
                    if (n) b->gcib[PAD(c, i, j, k)] = n->ileaf;
                }
            }
    /* unique donors in first-appearance order of the i-outer, k-inner scan */
    int *uniq = malloc(sizeof(int) * np), *cnt = malloc(sizeof(int) * np);
    int nu = 0;
    for (int i = 1 - ng; i <= ni + ng; i++)
        for (int j = 1 - ng; j <= nj + ng; j++)
            for (int k = 1 - ng; k <= nk + ng; k++) {
                int g = b->gcib[PAD(c, i, j, k)];
                if ((i < 1 || i > ni || j < 1 || j > nj || k < 1 || k > nk) && g != -777) {
                    int l;
                    for (l = 0; l < nu; l++) if (uniq[l] == g) { cnt[l]++; break; }
                    if (l == nu) { uniq[nu] = g; cnt[nu] = 1; nu++; }
                }
            }
    b->ntg = nu;
    b->tg = calloc((size_t)nu, sizeof(otarget));
    for (int u = 0; u < nu; u++) {
        otarget *t = &b->tg[u];
        t->donor = uniq[u]; t->ngc = cnt[u];
        t->ijk = malloc(sizeof(int32_t) * 3 * cnt[u]);
        t->xijk = malloc(sizeof(double) * 3 * cnt[u]);
        int n = 0;
        for (int i = 1 - ng; i <= ni + ng; i++)
            for (int j = 1 - ng; j <= nj + ng; j++)
                for (int k = 1 - ng; k <= nk + ng; k++)
                    if ((i < 1 || i > ni || j < 1 || j > nj || k < 1 || k > nk) && b->gcib[PAD(c, i, j, k)] == uniq[u]) {
                        t->ijk[n] = i; t->ijk[n + cnt[u]] = j; t->ijk[n + 2 * cnt[u]] = k;
                        t->xijk[n] = b->xi[i + ng - 1]; t->xijk[n + cnt[u]] = b->xj[j + ng - 1]; t->xijk[n + 2 * cnt[u]] = b->xk[k + ng - 1];
                        n++;
                    }
    }
    free(uniq); free(cnt);
}

/* local branches of ModCommunication_GetnGC_sources / SetGC_sources
 * (3_6/ModCommunication.f90:66-247) */
static void set_sources(orc_ctx *c) {
    int nb = c->nleaf;
    int *cnt = calloc((size_t)nb, sizeof(int));
    for (int ib = 0; ib < nb; ib++)
        for (int t = 0; t < c->b[ib].ntg; t++) cnt[c->b[ib].tg[t].donor - 1]++;
    for (int ib = 0; ib < nb; ib++) { c->b[ib].sc = calloc((size_t)(cnt[ib] ? cnt[ib] : 1), sizeof(osource)); c->b[ib].nsc = 0; }
    for (int ib = 0; ib < nb; ib++) {
        oblock *bt = &c->b[ib];
        for (int t = 0; t < bt->ntg; t++) {
            otarget *tg = &bt->tg[t];
            oblock *bs = &c->b[tg->donor - 1];
            osource *s = &bs->sc[bs->nsc++];
            s->recv = bt->iblock; s->recv_yin = bt->yin; s->ngc = tg->ngc;
            s->ijk = malloc(sizeof(int32_t) * 3 * tg->ngc);
            memcpy(s->ijk, tg->ijk, sizeof(int32_t) * 3 * tg->ngc);
            s->xijk = malloc(sizeof(double) * 3 * tg->ngc);
            if (s->recv_yin == bs->yin) {
                memcpy(s->xijk, tg->xijk, sizeof(double) * 3 * tg->ngc);
            } else {
                for (int g = 0; g < tg->ngc; g++) {
                    double in[3] = {tg->xijk[g], tg->xijk[g + tg->ngc], tg->xijk[g + 2 * tg->ngc]}, out[3];
                    coordconv(in, out);
                    s->xijk[g] = out[0]; s->xijk[g + tg->ngc] = out[1]; s->xijk[g + 2 * tg->ngc] = out[2];
                }
            }
        }
    }
    free(cnt);
}

int orc_finalize(orc_ctx *c) {
    c->nleaf = 0;
    label(c, &c->yin);
    c->nleaf_yin = c->nleaf;
    label(c, &c->yang);
    c->b = calloc((size_t)c->nleaf, sizeof(oblock));
    collect(c, &c->yin);
    collect(c, &c->yang);
    size_t np = (size_t)c->NI * c->NJ * c->NK, nint = (size_t)c->ni * c->nj * c->nk;
    for (int ib = 0; ib < c->nleaf; ib++) {
        oblock *b = &c->b[ib];
        init_grid(c, b);
        b->u = calloc(np * (ORC_NVAR + 1), sizeof(double));
        b->urk = calloc(np * (ORC_NVAR + 1), sizeof(double));
        b->r = calloc(nint * (ORC_NVAR + 1), sizeof(double));
        b->vw = calloc(nint, sizeof(double));
    }
    #pragma omp parallel for num_threads(g_threads) schedule(dynamic)
    for (int ib = 0; ib < c->nleaf; ib++) find_gc(c, &c->b[ib]);
    set_sources(c);
    return c->nleaf;
}

void orc_destroy(orc_ctx *c) {
    if (!c) return;
    for (int ib = 0; ib < c->nleaf; ib++) {
        oblock *b = &c->b[ib];
        free(b->xi); free(b->xj); free(b->xk); free(b->h); free(b->gcib);
        for (int t = 0; t < b->ntg; t++) { free(b->tg[t].ijk); free(b->tg[t].xijk); }
        for (int s = 0; s < b->nsc; s++) { free(b->sc[s].ijk); free(b->sc[s].xijk); }
        free(b->tg); free(b->sc);
        free(b->gamma1); free(b->gm1); free(b->rho0); free(b->g_over_rho0); free(b->p0_over_rho0);
        free(b->rho0T0); free(b->heat); free(b->xirsst);
        free(b->u); free(b->urk); free(b->r); free(b->vw);
    }
    free(c->b);
    node_free(&c->yin); node_free(&c->yang);
    free(c);
}

int orc_nblocks(const orc_ctx *c) { return c->nleaf; }
int orc_nleaf_yin(const orc_ctx *c) { return c->nleaf_yin; }

void orc_block_info(const orc_ctx *c, int ib, double range[6], int flags[3]) {
    const oblock *b = &c->b[ib];
    for (int d = 0; d < 3; d++) { range[d] = b->rng[d][0]; range[3 + d] = b->rng[d][1]; }
    flags[0] = b->yin; flags[1] = b->top; flags[2] = b->bottom;
}

void orc_block_centres(const orc_ctx *c, int ib, double *xi, double *xj, double *xk) {
    memcpy(xi, c->b[ib].xi, sizeof(double) * c->NI);
    memcpy(xj, c->b[ib].xj, sizeof(double) * c->NJ);
    memcpy(xk, c->b[ib].xk, sizeof(double) * c->NK);
}

void orc_gc_iblocks(const orc_ctx *c, int ib, int32_t *out) {
    size_t np = (size_t)c->NI * c->NJ * c->NK;
    for (size_t q = 0; q < np; q++) {
        /* interior cells keep the -1 of ModBlock_InitGrid only before FindGC;
         * ModGC_FindGC_Single resets everything to -777 first (3_6/ModGC.f90:34) */
        out[q] = c->b[ib].gcib[q];
    }
}

int orc_ntargets(const orc_ctx *c, int ib) { return c->b[ib].ntg; }
int orc_target_donor(const orc_ctx *c, int ib, int it) { return c->b[ib].tg[it].donor; }
int orc_target(const orc_ctx *c, int ib, int it, int32_t *ijk, double *xijk) {
    const otarget *t = &c->b[ib].tg[it];
    if (ijk) memcpy(ijk, t->ijk, sizeof(int32_t) * 3 * t->ngc);
    if (xijk) memcpy(xijk, t->xijk, sizeof(double) * 3 * t->ngc);
    return t->ngc;
}
int orc_nsources(const orc_ctx *c, int ib) { return c->b[ib].nsc; }
int orc_source(const orc_ctx *c, int ib, int is, int32_t *recv_block, int32_t *ijk, double *xijk) {
    const osource *s = &c->b[ib].sc[is];
    if (recv_block) *recv_block = s->recv;
    if (ijk) memcpy(ijk, s->ijk, sizeof(int32_t) * 3 * s->ngc);
    if (xijk) memcpy(xijk, s->xijk, sizeof(double) * 3 * s->ngc);
    return s->ngc;
}

/* ------------------------------------------------------------------ data */

/* Replication of the 1-D profiles into the per-cell arrays of
 * ModBlock_InitSSM (3_6/ModBlock.f90:252-264).  The oracle keeps them 1-D in r:
 * every (j,k) copy is identical. */
void orc_set_background(orc_ctx *c, int ib, const double *bg) {
    oblock *b = &c->b[ib];
    int NI = c->NI;
    double **arrs[8] = {&b->gamma1, &b->gm1, &b->rho0, &b->g_over_rho0, &b->p0_over_rho0, &b->rho0T0, &b->heat, &b->xirsst};
    for (int a = 0; a < 8; a++) { free(*arrs[a]); *arrs[a] = malloc(sizeof(double) * NI); }
    const double *g = bg, *rho0 = bg + NI, *p0 = bg + 2 * NI, *T0 = bg + 3 * NI, *g1 = bg + 4 * NI, *g3 = bg + 5 * NI,
                 *heat = bg + 6 * NI, *xi = bg + 7 * NI;
    for (int i = 0; i < NI; i++) {
        b->gamma1[i] = g1[i];
        b->gm1[i] = g3[i] - 1.0;
        b->rho0[i] = rho0[i];
        b->g_over_rho0[i] = g[i] / rho0[i];
        b->p0_over_rho0[i] = p0[i] / rho0[i];
        b->rho0T0[i] = rho0[i] * T0[i];
        b->heat[i] = heat[i] + 0.0;
        b->xirsst[i] = xi[i];
    }
}

void orc_set_state(orc_ctx *c, int ib, const double *u) {
    size_t n = (size_t)c->NI * c->NJ * c->NK * ORC_NVAR;
    memcpy(c->b[ib].u, u, sizeof(double) * n);
    memcpy(c->b[ib].urk, u, sizeof(double) * n);
}

void orc_get_state(const orc_ctx *c, int ib, double *u) {
    memcpy(u, c->b[ib].u, sizeof(double) * (size_t)c->NI * c->NJ * c->NK * ORC_NVAR);
}

/* ------------------------------------------------------------------ exchange */

/* ModMath_1D3D_interpolate_1D1D (3_6/ModMath.f90:250-293) for one point, all vars */
static void interp_point(const orc_ctx *c, const oblock *s, const double *f, const double x[3], double out[ORC_NVAR + 1]) {
    int ng = c->ng;
    double dxi = s->xi[ng + 1] - s->xi[ng], dxj = s->xj[ng + 1] - s->xj[ng], dxk = s->xk[ng + 1] - s->xk[ng];
    double base = (double)((float)(-ng) + 1.0f);
    double pi_ = base + (x[0] - s->xi[0]) / dxi;
    double pj = base + (x[1] - s->xj[0]) / dxj;
    double pk = base + (x[2] - s->xk[0]) / dxk;
    double ni_ = (double)((float)c->ni + 0.0f), nj_ = (double)((float)c->nj + 0.0f), nk_ = (double)((float)c->nk + 0.0f);
    pi_ = pi_ < ni_ ? pi_ : ni_; pi_ = pi_ > 1.0 ? pi_ : 1.0;
    pj = pj < nj_ ? pj : nj_; pj = pj > 1.0 ? pj : 1.0;
    pk = pk < nk_ ? pk : nk_; pk = pk > 1.0 ? pk : 1.0;
    int i0 = (int)floor(pi_), j0 = (int)floor(pj), k0 = (int)floor(pk);
    double wi = pi_ - (double)i0, wj = pj - (double)j0, wk = pk - (double)k0;
    size_t nv = (size_t)c->NI * c->NJ * c->NK;
    for (int v = 0; v < c->nv; v++) {
        const double *F = f + nv * v;
        double t = F[PAD(c, i0, j0, k0)] * (1.0 - wi) * (1.0 - wj) * (1.0 - wk);
        t = t + F[PAD(c, i0 + 1, j0, k0)] * (wi) * (1.0 - wj) * (1.0 - wk);
        t = t + F[PAD(c, i0, j0 + 1, k0)] * (1.0 - wi) * (wj) * (1.0 - wk);
        t = t + F[PAD(c, i0, j0, k0 + 1)] * (1.0 - wi) * (1.0 - wj) * (wk);
        t = t + F[PAD(c, i0 + 1, j0 + 1, k0)] * (wi) * (wj) * (1.0 - wk);
        t = t + F[PAD(c, i0, j0 + 1, k0 + 1)] * (1.0 - wi) * (wj) * (wk);
        t = t + F[PAD(c, i0 + 1, j0, k0 + 1)] * (wi) * (1.0 - wj) * (wk);
        t = t + F[PAD(c, i0 + 1, j0 + 1, k0 + 1)] * (wi) * (wj) * (wk);
        out[v] = t;
    }
}

/* ModYinYang_VecConv_1D (3_6/ModYinYang.f90:48-61) for one point; sin and
 * cos of the input phi as one sincos call (see coordconv) */
static void vecconv(const double rtp[3], double v[3]) {
    double o[3];
    coordconv(rtp, o);
    double sp, cp;
    sincos(rtp[2], &sp, &cp);
    double spo = sin(o[2]), sto = sin(o[1]);
    double v2 = -((sp * spo) * v[1]) - (cp / sto) * v[2];
    double v3 = -((sp * spo) * v[2]) + (cp / sto) * v[1];
    v[1] = v2; v[2] = v3;
}

/* ModCommunication_SendRecvGC_new (3_6/ModCommunication.f90:655-795), one rank.
 * Reads donors' interior (clamped trilinear touches donor ghosts only with an
 * exact-zero weight) and writes receivers' ghosts, so block order is free. */
void orc_exchange(orc_ctx *c, int which) {
    #pragma omp parallel for num_threads(g_threads) schedule(dynamic)
    for (int ib = 0; ib < c->nleaf; ib++) {
        oblock *s = &c->b[ib];
        const double *f = which ? s->urk : s->u;
        for (int is = 0; is < s->nsc; is++) {
            osource *sc = &s->sc[is];
            oblock *t = &c->b[sc->recv - 1];
            double *ft = which ? t->urk : t->u;
            size_t nv = (size_t)c->NI * c->NJ * c->NK;
            for (int g = 0; g < sc->ngc; g++) {
                double x[3] = {sc->xijk[g], sc->xijk[g + sc->ngc], sc->xijk[g + 2 * sc->ngc]};
                double val[ORC_NVAR + 1];
                interp_point(c, s, f, x, val);
                if (sc->recv_yin != s->yin) {
                    double bv[3] = {val[4], val[5], val[6]}, vv[3] = {val[1], val[2], val[3]};
                    vecconv(x, bv);
                    vecconv(x, vv);
                    val[4] = bv[0]; val[5] = bv[1]; val[6] = bv[2];
                    val[1] = vv[0]; val[2] = vv[1]; val[3] = vv[2];
                }
                size_t q = PAD(c, sc->ijk[g], sc->ijk[g + sc->ngc], sc->ijk[g + 2 * sc->ngc]);
                for (int v = 0; v < c->nv; v++) ft[q + nv * v] = val[v];
            }
        }
    }
}

/* ------------------------------------------------------------------ physics */

enum { RHO1 = 0, VR = 1, VT = 2, VP = 3, BR = 4, BT = 5, BP = 6, S1 = 7, PSI = 8 };

/* ModBoundary_Dynamo_primitives (3_6/ModBoundary.f90:10-55) */
static void boundary(const orc_ctx *c, const oblock *b, double *p) {
    int ni = c->ni, ng = c->ng;
    size_t nv = (size_t)c->NI * c->NJ * c->NK;
    static const double top_sign[8] = {1, -1, 1, 1, 1, -1, -1, 1};
    static const double bot_sign[8] = {1, -1, 1, 1, -1, 1, 1, 1};
    if (b->top)
        for (int i = ni + 1; i <= ni + ng; i++)
            for (int k = 1 - ng; k <= c->nk + ng; k++)
                for (int j = 1 - ng; j <= c->nj + ng; j++)
                    for (int v = 0; v < 8; v++) {
                        double x = p[nv * v + PAD(c, 2 * ni + 1 - i, j, k)];
                        p[nv * v + PAD(c, i, j, k)] = top_sign[v] < 0 ? -x : x;
                    }
    if (b->bottom)
        for (int i = 1 - ng; i <= 0; i++)
            for (int k = 1 - ng; k <= c->nk + ng; k++)
                for (int j = 1 - ng; j <= c->nj + ng; j++)
                    for (int v = 0; v < 8; v++) {
                        double x = p[nv * v + PAD(c, 1 - i, j, k)];
                        p[nv * v + PAD(c, i, j, k)] = bot_sign[v] < 0 ? -x : x;
                    }
}

/* ModDerivative_1st_O4_3D (3_6/ModDerivative.f90:5-34) on a padded array;
 * out is interior.  Coefficient (1./12./dx): float 1/12 then double divide. */
static void deriv(const orc_ctx *c, const double *f, double dx, int dir, double *out) {
    int ni = c->ni, nj = c->nj, nk = c->nk;
    double coef = (double)(1.0f / 12.0f) / dx;
    int di = dir == 1, dj = dir == 2, dk = dir == 3;
    for (int k = 1; k <= nk; k++)
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                double p2 = f[PAD(c, i + 2 * di, j + 2 * dj, k + 2 * dk)];
                double p1 = f[PAD(c, i + di, j + dj, k + dk)];
                double m1 = f[PAD(c, i - di, j - dj, k - dk)];
                double m2 = f[PAD(c, i - 2 * di, j - 2 * dj, k - 2 * dk)];
                out[INT(c, i, j, k)] = coef * ((((-p2) + 8.0 * p1) - 8.0 * m1) + m2);
            }
}

typedef struct { double *a[8]; } scratch8;

/* ModSpherical_A_dot_Grad_f (3_6/ModSpherical.f90:241-279):
 * A given by 3 padded arrays, f padded; out interior. */
static void a_dot_grad_f(const orc_ctx *c, const oblock *b, const double *A1, const double *A2, const double *A3,
                         const double *f, double *out, double *tmp, double *d) {
    int ni = c->ni, nj = c->nj, nk = c->nk, ng = c->ng;
    deriv(c, f, b->dxi, 1, d);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++)
        out[INT(c, i, j, k)] = A1[PAD(c, i, j, k)] * d[INT(c, i, j, k)];
    deriv(c, f, b->dxj, 2, d);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++)
        tmp[INT(c, i, j, k)] = A2[PAD(c, i, j, k)] * d[INT(c, i, j, k)];
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        double rinv = 1.0 / b->xi[i + ng - 1];
        out[INT(c, i, j, k)] = out[INT(c, i, j, k)] + tmp[INT(c, i, j, k)] * rinv;
    }
    deriv(c, f, b->dxk, 3, d);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        double rinv = 1.0 / b->xi[i + ng - 1];
        tmp[INT(c, i, j, k)] = (A3[PAD(c, i, j, k)] * d[INT(c, i, j, k)]) * rinv;
    }
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        double sinv = 1.0 / sin(b->xj[j + ng - 1]);
        out[INT(c, i, j, k)] = out[INT(c, i, j, k)] + sinv * tmp[INT(c, i, j, k)];
    }
}

/* ModSpherical_curl (3_6/ModSpherical.f90:38-102): A as 3 padded arrays */
static void curl(const orc_ctx *c, const oblock *b, const double *A1, const double *A2, const double *A3,
                 double *o1, double *o2, double *o3, double *w1, double *w2, double *d) {
    int ni = c->ni, nj = c->nj, nk = c->nk, ng = c->ng;
    size_t np = (size_t)c->NI * c->NJ * c->NK;
    /* r component: d_t(Ap sint) - d_p(At) */
    for (size_t q = 0; q < np; q++) {
        int j = (int)((q / c->NI) % c->NJ);
        w1[q] = A3[q] * sin(b->xj[j]);
    }
    deriv(c, w1, b->dxj, 2, o1);
    deriv(c, A2, b->dxk, 3, d);
    for (size_t q = 0; q < (size_t)ni * nj * nk; q++) o1[q] = o1[q] - d[q];
    /* t component: d_p(Ar/sint) - d_r(Ap r) */
    for (size_t q = 0; q < np; q++) {
        int i = (int)(q % c->NI), j = (int)((q / c->NI) % c->NJ);
        w1[q] = A1[q] / sin(b->xj[j]);
        w2[q] = A3[q] * b->xi[i];
    }
    deriv(c, w1, b->dxk, 3, o2);
    deriv(c, w2, b->dxi, 1, d);
    for (size_t q = 0; q < (size_t)ni * nj * nk; q++) o2[q] = o2[q] - d[q];
    /* p component: d_r(At r) - d_t(Ar) */
    for (size_t q = 0; q < np; q++) {
        int i = (int)(q % c->NI);
        w1[q] = A2[q] * b->xi[i];
    }
    deriv(c, w1, b->dxi, 1, o3);
    deriv(c, A1, b->dxj, 2, d);
    for (size_t q = 0; q < (size_t)ni * nj * nk; q++) o3[q] = o3[q] - d[q];
    /* rescale: all by 1/r, then r component by 1/sint */
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        double rinv = 1.0 / b->xi[i + ng - 1];
        size_t q = INT(c, i, j, k);
        o1[q] = o1[q] * rinv; o2[q] = o2[q] * rinv; o3[q] = o3[q] * rinv;
    }
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        double sinv = 1.0 / sin(b->xj[j + ng - 1]);
        size_t q = INT(c, i, j, k);
        o1[q] = o1[q] * sinv;
    }
}

/* ModEquation_Dynamo (3_6/ModEquation.f90:24-162) on padded primitive p */
static void equation(const orc_ctx *c, const oblock *b, const double *p, double *R) {
    int ni = c->ni, nj = c->nj, nk = c->nk, ng = c->ng;
    size_t np = (size_t)c->NI * c->NJ * c->NK, nn = (size_t)ni * nj * nk;
    const double *P[8];
    for (int v = 0; v < 8; v++) P[v] = p + np * v;
    double *p1 = malloc(sizeof(double) * np);
    double *w1 = malloc(sizeof(double) * np), *w2 = malloc(sizeof(double) * np), *w3 = malloc(sizeof(double) * np);
    double *t1 = malloc(sizeof(double) * nn), *t2 = malloc(sizeof(double) * nn), *t3 = malloc(sizeof(double) * nn);
    double *d = malloc(sizeof(double) * nn), *tmp = malloc(sizeof(double) * nn), *dd = malloc(sizeof(double) * nn);
    double *Rv[8];
    for (int v = 0; v < 8; v++) Rv[v] = R + nn * v;

    /* Get_p1 (:40-46) */
    for (size_t q = 0; q < np; q++) {
        int i = (int)(q % c->NI);
        p1[q] = b->gamma1[i] * b->p0_over_rho0[i] * P[RHO1][q] + b->gm1[i] * b->rho0T0[i] * P[S1][q];
    }
    /* Mass (:50-61): tmp = rho0*v; div */
    for (size_t q = 0; q < np; q++) {
        int i = (int)(q % c->NI), j = (int)((q / c->NI) % c->NJ);
        double r2 = b->xi[i] * b->xi[i];
        w1[q] = (b->rho0[i] * P[VR][q]) * r2;           /* A_scaled(:,:,:,1) = tmp1*r2 */
        w2[q] = (b->rho0[i] * P[VT][q]) * sin(b->xj[j]);  /* A_scaled(:,:,:,2) = tmp2*sint */
        w3[q] = b->rho0[i] * P[VP][q];                    /* A(:,:,:,3) */
    }
    /* ModSpherical_div (:104-159) */
    deriv(c, w1, b->dxi, 1, t1);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        double r = b->xi[i + ng - 1];
        t1[INT(c, i, j, k)] = t1[INT(c, i, j, k)] * (1.0 / (r * r));
    }
    deriv(c, w2, b->dxj, 2, t2);
    deriv(c, w3, b->dxk, 3, d);
    for (size_t q = 0; q < nn; q++) t2[q] = t2[q] + d[q];
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++)
        t2[INT(c, i, j, k)] = t2[INT(c, i, j, k)] * (1.0 / sin(b->xj[j + ng - 1]));
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++)
        t2[INT(c, i, j, k)] = t2[INT(c, i, j, k)] * (1.0 / b->xi[i + ng - 1]);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        size_t q = INT(c, i, j, k);
        double xr = b->xirsst[i + ng - 1];
        /* -1.0/Xi**2*div : (-1.0/(xi*xi))*div */
        Rv[RHO1][q] = (-1.0 / (xr * xr)) * (t1[q] + t2[q]);
    }

    /* Inertial (:66-74): -A_dot_nabla_B(v, v) (:281-343) */
    for (int comp = 0; comp < 3; comp++)
        a_dot_grad_f(c, b, P[VR], P[VT], P[VP], P[VR + comp], comp == 0 ? t1 : comp == 1 ? t2 : t3, tmp, d);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        size_t q = INT(c, i, j, k), pq = PAD(c, i, j, k);
        double rinv = 1.0 / b->xi[i + ng - 1];
        double xj = b->xj[j + ng - 1];
        /* cotant=cos(t)/sin(t) (ModSpherical.f90:296): one sincos call */
        double sxj, cxj;
        sincos(xj, &sxj, &cxj);
        double cot = cxj / sxj;
        double A2 = P[VT][pq], A3 = P[VP][pq], B1 = P[VR][pq], B2 = P[VT][pq], B3 = P[VP][pq];
        double tr = (-A2) * B2 - A3 * B3;
        t1[q] = t1[q] + tr * rinv;
        double tt = (-A3) * B3;
        tt = tt * cot;
        tt = tt + A2 * B1;
        t2[q] = t2[q] + tt * rinv;
        double tp = A3 * B2;
        tp = tp * cot;
        tp = tp + A3 * B1;
        t3[q] = t3[q] + tp * rinv;
        Rv[VR][q] = -t1[q]; Rv[VT][q] = -t2[q]; Rv[VP][q] = -t3[q];
    }

    /* Pressure gradient (:79-92): ModSpherical_Grad_f (:201-239) */
    deriv(c, p1, b->dxi, 1, t1);
    deriv(c, p1, b->dxj, 2, t2);
    deriv(c, p1, b->dxk, 3, t3);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        size_t q = INT(c, i, j, k);
        double rinv = 1.0 / b->xi[i + ng - 1];
        double sinv = 1.0 / sin(b->xj[j + ng - 1]);
        double g2 = t2[q] * rinv;
        double g3 = (t3[q] * rinv) * sinv;
        double rho0 = b->rho0[i + ng - 1];
        Rv[VR][q] = Rv[VR][q] - t1[q] / rho0;
        Rv[VT][q] = Rv[VT][q] - g2 / rho0;
        Rv[VP][q] = Rv[VP][q] - g3 / rho0;
    }

    /* Lorentz (:97-116): curl(B) x B / rho0 */
    curl(c, b, P[BR], P[BT], P[BP], t1, t2, t3, w1, w2, d);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        size_t q = INT(c, i, j, k), pq = PAD(c, i, j, k);
        double b1 = P[BR][pq], b2 = P[BT][pq], b3 = P[BP][pq];
        /* ModSpherical_cross (:26-36) */
        double c1 = t2[q] * b3 - t3[q] * b2;
        double c2 = t3[q] * b1 - t1[q] * b3;
        double c3 = t1[q] * b2 - t2[q] * b1;
        double rho0 = b->rho0[i + ng - 1];
        Rv[VR][q] = Rv[VR][q] + c1 / rho0;
        Rv[VT][q] = Rv[VT][q] + c2 / rho0;
        Rv[VP][q] = Rv[VP][q] + c3 / rho0;
    }

    /* Gravity (:121-127) */
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        size_t q = INT(c, i, j, k);
        Rv[VR][q] = Rv[VR][q] - b->g_over_rho0[i + ng - 1] * P[RHO1][PAD(c, i, j, k)];
    }

    /* Entropy advection (:132-140) and heating (:145-151) */
    a_dot_grad_f(c, b, P[VR], P[VT], P[VP], P[S1], t1, tmp, d);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        size_t q = INT(c, i, j, k);
        Rv[S1][q] = -t1[q];
        /* Artificial_heating_ratio*(total_heat)/rho0T0, ratio = 1.0 */
        Rv[S1][q] = Rv[S1][q] + (1.0 * b->heat[i + ng - 1]) / b->rho0T0[i + ng - 1];
    }

    /* Induction (:154-162): curl(v x B) over the padded range */
    {
        double *e1 = malloc(sizeof(double) * np), *e2 = malloc(sizeof(double) * np), *e3 = malloc(sizeof(double) * np);
        for (size_t q = 0; q < np; q++) {
            double a1 = P[VR][q], a2 = P[VT][q], a3 = P[VP][q], b1 = P[BR][q], b2 = P[BT][q], b3 = P[BP][q];
            e1[q] = a2 * b3 - a3 * b2;
            e2[q] = a3 * b1 - a1 * b3;
            e3[q] = a1 * b2 - a2 * b1;
        }
        curl(c, b, e1, e2, e3, Rv[BR], Rv[BT], Rv[BP], w1, w2, d);
        free(e1); free(e2); free(e3);
    }
    free(p1); free(w1); free(w2); free(w3); free(t1); free(t2); free(t3); free(d); free(tmp); free(dd);
}

/* ModLinReconstruct_minmod (3_6/ModLinReconstruct.f90:5-49) for one value triple */
static double minmod(double ul, double u, double ur) {
    double a = (ur - ul) * 0.5;
    double b = 2.0 * (ur - u);
    double c = 2.0 * (u - ul);
    double ab = a * b, bc = b * c;
    double m1 = ab > 0.0 ? ab : 0.0;   /* max(0.0,x) with x never NaN here */
    double m2 = bc > 0.0 ? bc : 0.0;
    double s = (m1 * m2) * 1.0e10 - (double)1.0e-30f;
    double half = copysign(0.5, s);
    double fa = fabs(a), fb = fabs(b), fc = fabs(c);
    double mn = fa < fb ? fa : fb;
    mn = mn < fc ? mn : fc;
    return (half + 0.5) * copysign(mn, a);
}

/* ModDiffusion_Aritificial_1 (3_6/ModDiffusion.f90:9-150), h integer, the
 * first nvar variables; MAX(0.0,NaN) pinned to 0.0. R accumulates (caller
 * zeroes it). */
static void diffusion(const orc_ctx *c, const oblock *b, const double *p, int h, double *R, int nvar) {
    int ni = c->ni, nj = c->nj, nk = c->nk, ng = c->ng;
    size_t np = (size_t)c->NI * c->NJ * c->NK, nn = (size_t)ni * nj * nk;
    double *cs = malloc(sizeof(double) * c->NI);
    for (int i = 0; i < c->NI; i++) cs[i] = (1.0 / b->xirsst[i]) * sqrt(b->gamma1[i] * b->p0_over_rho0[i]);
    int n[3] = {ni, nj, nk};
    for (int dir = 1; dir <= 3; dir++) {
        int di = dir == 1, dj = dir == 2, dk = dir == 3;
        const double *vel = p + np * dir;     /* primitive(:,:,:,direction1+1) */
        int nf = n[dir - 1] + 1;
        (void)nf;
        for (int v = 0; v < nvar; v++) {
            const double *u = p + np * v;
            for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
                double fl[2];
                for (int side = 0; side < 2; side++) {
                    /* face between cell (x-1) and cell x, where x = (i,j,k)+side*dir */
                    int xi_ = i + side * di, xj_ = j + side * dj, xk_ = k + side * dk;
                    int yi = xi_ - di, yj = xj_ - dj, yk = xk_ - dk;
                    double ux = u[PAD(c, xi_, xj_, xk_)], uy = u[PAD(c, yi, yj, yk)];
                    double dx_ = minmod(u[PAD(c, xi_ - di, xj_ - dj, xk_ - dk)], ux, u[PAD(c, xi_ + di, xj_ + dj, xk_ + dk)]);
                    double dy_ = minmod(u[PAD(c, yi - di, yj - dj, yk - dk)], uy, u[PAD(c, yi + di, yj + dj, yk + dk)]);
                    double num = ((ux - dx_ * 0.5) - uy) - dy_ * 0.5;
                    double ph = 1.0 + (double)h * (num / (ux - uy) - 1.0);
                    ph = ph > 0.0 ? ph : 0.0;   /* NaN -> 0.0 */
                    double f = ((-0.5) * ph) * num;
                    /* c(1:ni+1,:,:) etc. span the padded transverse range while
                     * flux spans 1:nj,1:nk (:63,96,130); the non-conforming section
                     * is read from its own lower bound, i.e. at transverse index
                     * (index - ng).  Reproduced as the reference computes it. */
                    int ci = di ? xi_ : i - ng, cj = dj ? xj_ : j - ng, ck = dk ? xk_ : k - ng;
                    double cx = fabs(vel[PAD(c, ci, cj, ck)]) + cs[ci + ng - 1];
                    double cy = fabs(vel[PAD(c, ci - di, cj - dj, ck - dk)]) + cs[ci - di + ng - 1];
                    fl[side] = (f * (cx + cy)) * 0.5;
                }
                size_t q = INT(c, i, j, k) + nn * v;
                double L;
                if (dir == 1) L = b->dxi;
                else if (dir == 2) L = b->dxj * b->xi[i + ng - 1];
                else L = (b->dxk * b->xi[i + ng - 1]) * sin(b->xj[j + ng - 1]);
                R[q] = R[q] + (fl[0] - fl[1]) / L;
            }
        }
    }
    free(cs);
}

void orc_rhs(orc_ctx *c, int ib, double *r_out) {
    oblock *b = &c->b[ib];
    boundary(c, b, b->u);
    equation(c, b, b->u, r_out);
}

void orc_diffusion(orc_ctx *c, int ib, double *r_out) {
    oblock *b = &c->b[ib];
    memset(r_out, 0, sizeof(double) * (size_t)c->ni * c->nj * c->nk * 8);
    diffusion(c, b, b->u, 2, r_out, 8);
}

/* ModWaveSpeed_Dynamo + ModTimeStep_Dynamo (3_6/ModWaveSpeed.f90:8-27,
 * 3_6/ModTimeStep.f90:10-30) */
double orc_dt(orc_ctx *c, double cfl) {
    double dt = (double)1.0e20f;   /* dt=1.e20 (default real) */
    int ni = c->ni, nj = c->nj, nk = c->nk, ng = c->ng;
    for (int ib = 0; ib < c->nleaf; ib++) {
        oblock *b = &c->b[ib];
        size_t np = (size_t)c->NI * c->NJ * c->NK;
        double mn = INFINITY;
        int first = 1;
        for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
            size_t q = PAD(c, i, j, k);
            int ii = i + ng - 1;
            double vr = b->u[q + np * VR], vt = b->u[q + np * VT], vp = b->u[q + np * VP];
            double br = b->u[q + np * BR], bt = b->u[q + np * BT], bp = b->u[q + np * BP];
            double w = (1.0 / b->xirsst[ii]) * sqrt(b->gamma1[ii] * b->p0_over_rho0[ii]) + sqrt(vr * vr + vt * vt + vp * vp);
            w = w + sqrt(br * br + bt * bt + bp * bp) / sqrt(b->rho0[ii]);
            b->vw[INT(c, i, j, k)] = w;
            double x = b->h[INT(c, i, j, k)] / w;
            if (first || x < mn) { mn = x; first = 0; }
        }
        double dt1 = cfl * mn;
        dt = dt < dt1 ? dt : dt1;
    }
    return dt;
}

/* ModDivB_GLM (3_6/ModDivB.f90:13-38) on the padded stage input p, after
 * ModEquation_Dynamo: R(B) = R(B) - grad(psi) and R(psi) = -c_h^2 div(B) -
 * (c_h / h) psi, c_h = v_wave of the step (ModWaveSpeed at the step's start),
 * h = h_LLL.  div is ModSpherical_div (3_6/ModSpherical.f90:104-159), grad
 * ModSpherical_Grad_f (:201-239), each in the reference's operation order. */
static void glm(const orc_ctx *c, const oblock *b, const double *p, double *R) {
    int ni = c->ni, nj = c->nj, nk = c->nk, ng = c->ng;
    size_t np = (size_t)c->NI * c->NJ * c->NK, nn = (size_t)ni * nj * nk;
    const double *Br = p + np * BR, *Bt = p + np * BT, *Bp = p + np * BP, *psi = p + np * PSI;
    double *w = malloc(sizeof(double) * np);
    double *dv = malloc(sizeof(double) * nn), *d23 = malloc(sizeof(double) * nn), *d = malloc(sizeof(double) * nn);
    double *g1 = malloc(sizeof(double) * nn), *g2 = malloc(sizeof(double) * nn), *g3 = malloc(sizeof(double) * nn);
    /* div: A_scaled(1) = A_r r**2, A_scaled(2) = A_t sin(t) */
    for (size_t q = 0; q < np; q++) {
        int i = (int)(q % c->NI);
        w[q] = Br[q] * (b->xi[i] * b->xi[i]);
    }
    deriv(c, w, b->dxi, 1, dv);
    for (size_t q = 0; q < np; q++) {
        int j = (int)((q / c->NI) % c->NJ);
        w[q] = Bt[q] * sin(b->xj[j]);
    }
    deriv(c, w, b->dxj, 2, d23);
    deriv(c, Bp, b->dxk, 3, d);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        size_t q = INT(c, i, j, k);
        double r = b->xi[i + ng - 1];
        double r2_inverse = 1.0 / (r * r), r_inverse = 1.0 / r, sint_inverse = 1.0 / sin(b->xj[j + ng - 1]);
        double a = dv[q] * r2_inverse;
        double t = d23[q] + d[q];
        t = t * sint_inverse;
        t = t * r_inverse;
        dv[q] = a + t;
    }
    /* grad psi */
    deriv(c, psi, b->dxi, 1, g1);
    deriv(c, psi, b->dxj, 2, g2);
    deriv(c, psi, b->dxk, 3, g3);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        size_t q = INT(c, i, j, k);
        double r_inverse = 1.0 / b->xi[i + ng - 1], sint_inverse = 1.0 / sin(b->xj[j + ng - 1]);
        g2[q] = g2[q] * r_inverse;
        g3[q] = g3[q] * r_inverse;
        g3[q] = g3[q] * sint_inverse;
        R[q + nn * BR] = R[q + nn * BR] - g1[q];
        R[q + nn * BT] = R[q + nn * BT] - g2[q];
        R[q + nn * BP] = R[q + nn * BP] - g3[q];
        double vw = b->vw[q];
        R[q + nn * PSI] = -((vw * vw) * dv[q]) - (vw / b->h[q]) * psi[PAD(c, i, j, k)];
    }
    free(w); free(dv); free(d23); free(d); free(g1); free(g2); free(g3);
}

/* ModAdvance_rk4 (3_6/ModAdvance.f90:18-112), one rank */
double orc_step(orc_ctx *c, double cfl) {
    double dt = orc_dt(c, cfl);
    int ni = c->ni, nj = c->nj, nk = c->nk;
    size_t np = (size_t)c->NI * c->NJ * c->NK, nn = (size_t)ni * nj * nk;
    for (int rk = 1; rk <= 4; rk++) {
        #pragma omp parallel for num_threads(g_threads) schedule(dynamic)
        for (int ib = 0; ib < c->nleaf; ib++) {
            oblock *b = &c->b[ib];
            double *in = rk > 1 ? b->urk : b->u;
            if (c->lin_on) {
                for (int v = 0; v < 8; v++)
                    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++)
                        b->r[INT(c, i, j, k) + nn * v] = c->lin_lambda * in[PAD(c, i, j, k) + np * v];
            } else {
                boundary(c, b, in);
                equation(c, b, in, b->r);
                if (c->nv > 8) glm(c, b, in, b->r);
            }
            double div = (double)(5.0f - (float)rk);
            double *out = rk < 4 ? b->urk : b->u;
            for (int v = 0; v < c->nv; v++)
                for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
                    size_t q = PAD(c, i, j, k) + np * v;
                    out[q] = b->u[q] + (dt * b->r[INT(c, i, j, k) + nn * v]) / div;
                }
        }
        orc_exchange(c, rk < 4 ? 1 : 0);
    }
    #pragma omp parallel for num_threads(g_threads) schedule(dynamic)
    for (int ib = 0; ib < c->nleaf; ib++) {
        oblock *b = &c->b[ib];
        memset(b->r, 0, sizeof(double) * nn * c->nv);
        if (c->lin_on) {
            for (int v = 0; v < 8; v++)
                for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++)
                    b->r[INT(c, i, j, k) + nn * v] = c->lin_mu * b->u[PAD(c, i, j, k) + np * v];
        } else {
            diffusion(c, b, b->u, 2, b->r, c->nv);
        }
        for (int v = 0; v < c->nv; v++)
            for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
                size_t q = PAD(c, i, j, k) + np * v;
                b->u[q] = b->u[q] + (0.5 * dt) * b->r[INT(c, i, j, k) + nn * v];
            }
    }
    orc_exchange(c, 0);
    return dt;
}

void orc_set_glm(orc_ctx *c, int on) { c->nv = on ? 9 : 8; }

void orc_set_psi(orc_ctx *c, int ib, const double *psi) {
    size_t np = (size_t)c->NI * c->NJ * c->NK;
    memcpy(c->b[ib].u + np * PSI, psi, sizeof(double) * np);
    memcpy(c->b[ib].urk + np * PSI, psi, sizeof(double) * np);
}

void orc_get_psi(const orc_ctx *c, int ib, double *psi) {
    size_t np = (size_t)c->NI * c->NJ * c->NK;
    memcpy(psi, c->b[ib].u + np * PSI, sizeof(double) * np);
}

void orc_set_linear_rhs(orc_ctx *c, int on, double lambda, double mu) {
    c->lin_on = on;
    c->lin_lambda = lambda;
    c->lin_mu = mu;
}

/* Yin-Yang point map and the VecConv coefficients (sin p sin p', cos p / sin t')
 * for one donor-frame point, as used by vecconv() above. */
void orc_coordconv(const double in[3], double out[3]) { coordconv(in, out); }
void orc_vecconv_coeffs(const double x[3], double out[2]) {
    double o[3];
    coordconv(x, o);
    double sp, cp;
    sincos(x[2], &sp, &cp);
    double spo = sin(o[2]), sto = sin(o[1]);
    out[0] = sp * spo;
    out[1] = cp / sto;
}

/* ======================================================================
 * Heat conduction oracle.  The reference has no heat operator
 * (3_6/ModHeatConduction.f90:15-16); the method restated here is the one the
 * product defines: backward Euler (V/dt)T + sum_f kappa_f S_f/D_f (T - T_nb)
 * = (V/dt) T^n with face areas/volumes/distances of ModMultigrid_InitGrid
 * (3_5/ModMultigrid.f90:150-233) on the block's uniform phi spacing
 * dphi = (phi_hi - phi_lo)/m (so every coefficient depends on r and theta
 * only), the finite-volume assembly of ModPFSS_setup
 * (3_5/ModPFSS.f90:68-84), insulated radial boundaries, kappa(r) =
 * kappa0 (r/r_ref)^-3; per-block multigrid levels with face ghosts filled by
 * trilinear inter/extrapolation inside the donor found by FindNode (exact for
 * fields linear in r, theta, phi, whatever the two blocks' spacings; weights
 * within 1e-12 of 0 or 1 snapped to 0 or 1); V-cycle
 * with damped Jacobi (omega 0.8, 2+2 sweeps, 2 on the 1-cell level), summed
 * restriction, injection prolongation; flexible CG (FCG(1)).  Serial FP64.
 * Arrays: Fortran order (i fastest), padded (m+2)^3 with one ghost layer.
 * ====================================================================== */

typedef struct {
    int m;
    double *V, *Ai, *Aj, *Ak, *off;    /* per block: m^3, (m+1)m m (i fastest, face index first) ... */
    int ng;                             /* ghost entries */
    long *gdst, *gsrc;                  /* global padded indices (block*(m+2)^3 + cell) */
    double *gw;
} hlev;

struct orc_heat {
    orc_ctx *c;
    int nlev, nb;
    hlev L[8];
    double dt_expl;
    double coarse_tol;      /* coarsest-level CG: stop at r.z <= tol^2 (r.z)_0 */
    int coarse_max;         /* ... or after this many iterations */
};

/* Coarsest-level CG defaults (the product's, heat.hip) */
#define HEAT_COARSE_TOL 1.0e-2
#define HEAT_COARSE_MAX 50

#define HP(m, i, j, k) ((long)(i) + (long)((m) + 2) * ((long)(j) + (long)((m) + 2) * (long)(k)))  /* padded, 0..m+1 */

static void lev_centres(const oblock *b, int m, double *xc[3], double *xf[3]) {
    for (int a = 0; a < 3; a++)
        for (int i = 0; i <= m + 1; i++) {
            xc[a][i] = (b->rng[a][0] * (double)((float)(m - i) + 0.5f) + b->rng[a][1] * (double)((float)i - 0.5f)) / (double)m;
            if (i >= 1) xf[a][i - 1] = (b->rng[a][0] * (double)(m + 1 - i) + b->rng[a][1] * (double)(i - 1)) / (double)m;
        }
}

/* kappa(r) = kappa0 (r/r_ref)^kpow, kpow -3 (heat conduction) or 0 (uniform:
 * the PFSS Laplacian of 3_5/ModPFSS.f90:68-84) */
static double h_kappa(double r, double kappa0, double r_ref, int kpow) {
    if (kpow == 0) return kappa0;
    double x = r / r_ref;
    return kappa0 / (x * x * x);
}

orc_heat *orc_heat_create(orc_ctx *c, double kappa0, double r_ref) { return orc_heat_create_opt(c, kappa0, r_ref, -3, 0); }

/* top_dirichlet: the top radial face holds the value 0 (3_5/ModPFSS.f90:88-91:
 * Aii = Aii - Ai(top), Ai(top) = 0), i.e. its coefficient enters the diagonal
 * twice and couples to nothing; otherwise the top is insulated */
orc_heat *orc_heat_create_opt(orc_ctx *c, double kappa0, double r_ref, int kpow, int top_dirichlet) {
    orc_heat *H = calloc(1, sizeof(orc_heat));
    H->c = c; H->nb = c->nleaf;
    int n = c->ni;
    int l = 0;
    double worst = 0.0;
    for (int m = n; m >= 1; m /= 2, l++) {
        hlev *L = &H->L[l];
        L->m = m;
        long M3 = (long)m * m * m, P = (long)(m + 2) * (m + 2) * (m + 2);
        L->V = calloc((size_t)H->nb * M3, sizeof(double));
        L->off = calloc((size_t)H->nb * M3, sizeof(double));
        L->Ai = calloc((size_t)H->nb * (m + 1) * m * m, sizeof(double));
        L->Aj = calloc((size_t)H->nb * (m + 1) * m * m, sizeof(double));
        L->Ak = calloc((size_t)H->nb * (m + 1) * m * m, sizeof(double));
        L->gdst = malloc(sizeof(long) * (size_t)H->nb * 6 * m * m);
        L->gsrc = malloc(sizeof(long) * (size_t)H->nb * 6 * m * m);
        L->gw = malloc(sizeof(double) * 3 * (size_t)H->nb * 6 * m * m);
        L->ng = 0;
        double *xc[3], *xf[3], *dc[3], *df[3];
        for (int a = 0; a < 3; a++) {
            xc[a] = malloc(sizeof(double) * (m + 2)); xf[a] = malloc(sizeof(double) * (m + 1));
            dc[a] = malloc(sizeof(double) * (m + 2)); df[a] = malloc(sizeof(double) * (m + 1));
        }
        for (int ib = 0; ib < H->nb; ib++) {
            oblock *b = &c->b[ib];
            lev_centres(b, m, xc, xf);
            double dx[3];
            for (int a = 0; a < 3; a++) dx[a] = (b->rng[a][1] - b->rng[a][0]) / (double)m;
            double *V = L->V + (size_t)ib * M3, *off = L->off + (size_t)ib * M3;
            double *Ai = L->Ai + (size_t)ib * (m + 1) * m * m, *Aj = L->Aj + (size_t)ib * (m + 1) * m * m,
                   *Ak = L->Ak + (size_t)ib * (m + 1) * m * m;
            /* face index along its own axis first: Ai[(fi*m + j)*m + k], Aj[(i*(m+1)+fj)*m + k], Ak[(i*m+j)*(m+1)+fk] */
            for (int i = 0; i <= m; i++) for (int j = 0; j < m; j++) for (int k = 0; k < m; k++) {
                double kap = h_kappa(xf[0][i], kappa0, r_ref, kpow);
                double S = xf[0][i] * xf[0][i] * (cos(xf[1][j]) - cos(xf[1][j + 1])) * dx[2];
                double cf = kap * S / dx[0];
                if ((i == 0 && b->bottom) || (i == m && b->top && !top_dirichlet)) cf = 0.0;
                Ai[((long)i * m + j) * m + k] = cf;
            }
            for (int i = 0; i < m; i++) for (int j = 0; j <= m; j++) for (int k = 0; k < m; k++) {
                double kap = h_kappa(xc[0][i + 1], kappa0, r_ref, kpow);
                double S = 0.5 * (xf[0][i + 1] * xf[0][i + 1] - xf[0][i] * xf[0][i]) * dx[2] * sin(xf[1][j]);
                Aj[((long)i * (m + 1) + j) * m + k] = kap * S / (xc[0][i + 1] * dx[1]);
            }
            for (int i = 0; i < m; i++) for (int j = 0; j < m; j++) for (int k = 0; k <= m; k++) {
                double kap = h_kappa(xc[0][i + 1], kappa0, r_ref, kpow);
                double S = 0.5 * (xf[0][i + 1] * xf[0][i + 1] - xf[0][i] * xf[0][i]) * (xf[1][j + 1] - xf[1][j]);
                Ak[((long)i * m + j) * (m + 1) + k] = kap * S / (xc[0][i + 1] * dx[2] * sin(xc[1][j + 1]));
            }
            for (int i = 0; i < m; i++) for (int j = 0; j < m; j++) for (int k = 0; k < m; k++) {
                long q = (long)i + (long)m * ((long)j + (long)m * k);
                V[q] = 1.0 / 3.0 * (xf[0][i + 1] * xf[0][i + 1] * xf[0][i + 1] - xf[0][i] * xf[0][i] * xf[0][i]) *
                       (cos(xf[1][j]) - cos(xf[1][j + 1])) * dx[2];
                off[q] = Ai[((long)i * m + j) * m + k] + Ai[((long)(i + 1) * m + j) * m + k] +
                         Aj[((long)i * (m + 1) + j) * m + k] + Aj[((long)i * (m + 1) + j + 1) * m + k] +
                         Ak[((long)i * m + j) * (m + 1) + k] + Ak[((long)i * m + j) * (m + 1) + k + 1];
                if (top_dirichlet && b->top && i == m - 1) off[q] = off[q] + Ai[((long)m * m + j) * m + k];
                if (l == 0) {
                    double r = xc[0][i + 1], th = xc[1][j + 1];
                    double kap = h_kappa(r, kappa0, r_ref, kpow);
                    double h2 = r * dx[1], h3 = r * sin(th) * dx[2];
                    double v = 2.0 * kap * (1.0 / (dx[0] * dx[0]) + 1.0 / (h2 * h2) + 1.0 / (h3 * h3));
                    if (v > worst) worst = v;
                }
            }
            /* face ghosts */
            for (int I = 0; I <= m + 1; I++) for (int J = 0; J <= m + 1; J++) for (int K = 0; K <= m + 1; K++) {
                int out = (I == 0 || I == m + 1) + (J == 0 || J == m + 1) + (K == 0 || K == m + 1);
                if (out != 1) continue;
                double x[3] = {xc[0][I], xc[1][J], xc[2][K]};
                const onode *nd = find_node(c, x, b->yin);
                if (!nd) continue;
                int d = nd->ileaf - 1;
                oblock *db = &c->b[d];
                lev_centres(db, m, dc, df);
                if (db->yin != b->yin) { double o[3]; coordconv(x, o); x[0] = o[0]; x[1] = o[1]; x[2] = o[2]; }
                /* linear inter/extrapolation inside the donor block: p in [0.5, m+0.5],
                 * stencil base in [1, m-1] (interior cells only); m = 1 copies the cell */
                int base[3]; double w[3];
                for (int a = 0; a < 3; a++) {
                    if (m == 1) { base[a] = 1; w[a] = 0.0; continue; }
                    double dxa = dc[a][2] - dc[a][1];
                    double p = (x[a] - dc[a][0]) / dxa;
                    p = p < (double)m + 0.5 ? p : (double)m + 0.5;
                    p = p > 0.5 ? p : 0.5;
                    int bs = (int)floor(p);
                    bs = bs < 1 ? 1 : (bs > m - 1 ? m - 1 : bs);
                    base[a] = bs;
                    w[a] = p - (double)bs;
                    if (fabs(w[a]) < 1e-12) w[a] = 0.0;
                    else if (fabs(w[a] - 1.0) < 1e-12) w[a] = 1.0;
                }
                L->gdst[L->ng] = (long)ib * P + HP(m, I, J, K);
                L->gsrc[L->ng] = (long)d * P + HP(m, base[0], base[1], base[2]);
                L->gw[3 * L->ng] = w[0]; L->gw[3 * L->ng + 1] = w[1]; L->gw[3 * L->ng + 2] = w[2];
                L->ng++;
            }
        }
        for (int a = 0; a < 3; a++) { free(xc[a]); free(xf[a]); free(dc[a]); free(df[a]); }
    }
    H->nlev = l;
    H->dt_expl = 1.0 / worst;
    H->coarse_tol = HEAT_COARSE_TOL;
    H->coarse_max = HEAT_COARSE_MAX;
    return H;
}

void orc_heat_destroy(orc_heat *H) {
    if (!H) return;
    for (int l = 0; l < H->nlev; l++) {
        hlev *L = &H->L[l];
        free(L->V); free(L->off); free(L->Ai); free(L->Aj); free(L->Ak); free(L->gdst); free(L->gsrc); free(L->gw);
    }
    free(H);
}

double orc_heat_dt_explicit(const orc_heat *H) { return H->dt_expl; }
int orc_heat_levels(const orc_heat *H) { return H->nlev; }

static void h_exchange(const hlev *L, double *x) {
    const int m = L->m;
    const long SJ = m + 2, SK = (long)(m + 2) * (m + 2);
    for (int e = 0; e < L->ng; e++) {
        const double *F = x + L->gsrc[e];
        double wi = L->gw[3 * e], wj = L->gw[3 * e + 1], wk = L->gw[3 * e + 2];
        double t = F[0] * (1.0 - wi) * (1.0 - wj) * (1.0 - wk);
        t = t + F[1] * (wi) * (1.0 - wj) * (1.0 - wk);
        t = t + F[SJ] * (1.0 - wi) * (wj) * (1.0 - wk);
        t = t + F[SK] * (1.0 - wi) * (1.0 - wj) * (wk);
        t = t + F[1 + SJ] * (wi) * (wj) * (1.0 - wk);
        t = t + F[SJ + SK] * (1.0 - wi) * (wj) * (wk);
        t = t + F[1 + SK] * (wi) * (1.0 - wj) * (wk);
        t = t + F[1 + SJ + SK] * (wi) * (wj) * (wk);
        x[L->gdst[e]] = t;
    }
}

/* mode 0: y = A x; 1: y = r - A x; 2: y(padded) = x + omega (r - A x)/diag */
static void h_apply(const hlev *L, int nb, double inv_dt, const double *x, const double *r, double *y, int mode, double omega) {
    const int m = L->m;
    const long M3 = (long)m * m * m, P = (long)(m + 2) * (m + 2) * (m + 2);
    for (int b = 0; b < nb; b++) {
        const double *X = x + (long)b * P;
        const double *Ai = L->Ai + (size_t)b * (m + 1) * m * m, *Aj = L->Aj + (size_t)b * (m + 1) * m * m,
                     *Ak = L->Ak + (size_t)b * (m + 1) * m * m;
        for (int k = 0; k < m; k++) for (int j = 0; j < m; j++) for (int i = 0; i < m; i++) {
            long q = (long)i + (long)m * ((long)j + (long)m * k);
            long pc = HP(m, i + 1, j + 1, k + 1);
            double cim = Ai[((long)i * m + j) * m + k], cip = Ai[((long)(i + 1) * m + j) * m + k];
            double cjm = Aj[((long)i * (m + 1) + j) * m + k], cjp = Aj[((long)i * (m + 1) + j + 1) * m + k];
            double ckm = Ak[((long)i * m + j) * (m + 1) + k], ckp = Ak[((long)i * m + j) * (m + 1) + k + 1];
            double diag = L->V[(long)b * M3 + q] * inv_dt + L->off[(long)b * M3 + q];
            double nbv = cim * X[pc - 1] + cip * X[pc + 1];
            nbv = nbv + cjm * X[pc - (m + 2)] + cjp * X[pc + (m + 2)];
            nbv = nbv + ckm * X[pc - (long)(m + 2) * (m + 2)] + ckp * X[pc + (long)(m + 2) * (m + 2)];
            double ax = diag * X[pc] - nbv;
            long io = (long)b * M3 + q;
            if (mode == 0) y[io] = ax;
            else if (mode == 1) y[io] = r[io] - ax;
            else y[(long)b * P + pc] = X[pc] + omega * (r[io] - ax) / diag;
        }
    }
}

typedef struct { double *Z, *Z2, *R, *RES; } hwork;
static void h_pad(int m, int nb, const double *in, double *out);

/* Sum of per-row values in the order of the product's one-workgroup reduction
 * (heat.hip, k_hcoarse_cg): 1024 threads, thread t adds rows t, t + 1024, ...
 * from 0.0; each 64-lane wavefront reduces by halving shifts (lane l takes
 * l + 32, then l + 16, ...); the 16 wavefront sums are added in order. */
static double wg_sum(const double *v, long n) {
    double acc[1024], red[16];
    for (int t = 0; t < 1024; t++) {
        double a = 0.0;
        for (long q = t; q < n; q += 1024) a += v[q];
        acc[t] = a;
    }
    for (int w = 0; w < 16; w++) {
        double *b = acc + 64 * w;
        for (int off = 32; off > 0; off >>= 1)
            for (int l = 0; l < off; l++) b[l] = b[l] + b[l + off];
        red[w] = b[0];
    }
    double t = red[0];
    for (int w = 1; w < 16; w++) t = t + red[w];
    return t;
}

/* Coarsest level (one cell per block): Jacobi-preconditioned CG from x = 0,
 * ModLinearSolver_JCB_CG (3_5/ModLinearSolver.f90:86-180) as a method, with
 * the Jacobi diagonal as preconditioner and a relative stop on r.z. */
static void h_coarse_cg(orc_heat *H, hwork *W, int l, double inv_dt) {
    const hlev *L = &H->L[l];
    const int nb = H->nb;
    const long P = 27;          /* (m + 2)^3 at m = 1 */
    const long pc = HP(1, 1, 1, 1);
    double *x = W[l].Z;
    double *r = malloc(sizeof(double) * nb), *pp = calloc((size_t)nb * P, sizeof(double));
    double *Ap = malloc(sizeof(double) * nb), *dg = malloc(sizeof(double) * nb), *t = malloc(sizeof(double) * nb);
    memset(x, 0, sizeof(double) * (size_t)nb * P);
    for (int b = 0; b < nb; b++) {
        dg[b] = L->V[b] * inv_dt + L->off[b];
        r[b] = W[l].R[b];
        const double z = r[b] / dg[b];
        pp[(long)b * P + pc] = z;
        t[b] = r[b] * z;
    }
    double rz = wg_sum(t, nb);
    const double stop = (H->coarse_tol * H->coarse_tol) * rz;
    for (int it = 0; it < H->coarse_max && rz > stop; it++) {
        h_exchange(L, pp);
        h_apply(L, nb, inv_dt, pp, NULL, Ap, 0, 0.0);
        for (int b = 0; b < nb; b++) t[b] = pp[(long)b * P + pc] * Ap[b];
        const double pAp = wg_sum(t, nb);
        if (!(pAp > 0.0)) break;
        const double alpha = rz / pAp;
        for (int b = 0; b < nb; b++) {
            x[(long)b * P + pc] = x[(long)b * P + pc] + alpha * pp[(long)b * P + pc];
            r[b] = r[b] - alpha * Ap[b];
            t[b] = r[b] * (r[b] / dg[b]);
        }
        const double rzn = wg_sum(t, nb);
        const double beta = rzn / rz;
        rz = rzn;
        for (int b = 0; b < nb; b++) pp[(long)b * P + pc] = r[b] / dg[b] + beta * pp[(long)b * P + pc];
    }
    free(r); free(pp); free(Ap); free(dg); free(t);
}

/* Linear prolongation, stored: every fine cell adds zc + 1/4 of the coarse
 * differences toward its three face neighbours on its side of the parent
 * (cell-centred linear interpolation from face neighbours only; exact for
 * fields linear in the block's index space).  Coarse ghosts come from the
 * coarse level's face-ghost fill; at an insulated radial face (no donor) the
 * neighbour is the parent itself (zero gradient). */
static void h_prolong_linear(const orc_heat *H, int l, const double *Zc_all, double *Zf_all) {
    const int nb = H->nb, m = H->L[l].m, mc = m / 2;
    const long P = (long)(m + 2) * (m + 2) * (m + 2), Pc = (long)(mc + 2) * (mc + 2) * (mc + 2);
    for (int b = 0; b < nb; b++) {
        const oblock *ob = &H->c->b[b];
        const double *Zc = Zc_all + (long)b * Pc;
        for (int k = 0; k < m; k++) for (int j = 0; j < m; j++) for (int i = 0; i < m; i++) {
            const int ic = i / 2 + 1, jc = j / 2 + 1, kc = k / 2 + 1;
            const int si = (i & 1) ? 1 : -1, sj = (j & 1) ? 1 : -1, sk = (k & 1) ? 1 : -1;
            const double z0 = Zc[HP(mc, ic, jc, kc)];
            const int in = ic + si;
            double zi = Zc[HP(mc, in, jc, kc)];
            if ((in == 0 && ob->bottom) || (in == mc + 1 && ob->top)) zi = z0;
            const double zj = Zc[HP(mc, ic, jc + sj, kc)], zk = Zc[HP(mc, ic, jc, kc + sk)];
            double v = z0 + 0.25 * (zi - z0);
            v = v + 0.25 * (zj - z0);
            v = v + 0.25 * (zk - z0);
            double *zf = &Zf_all[(long)b * P + HP(m, i + 1, j + 1, k + 1)];
            *zf = *zf + v;
        }
    }
}

static void h_vcycle(orc_heat *H, hwork *W, double inv_dt, const double *r0, double *z0_out) {
    const int nb = H->nb;
    W[0].R = (double *)r0;
    for (int l = 0; l < H->nlev; l++) {
        const hlev *L = &H->L[l];
        const int m = L->m;
        const long P = (long)(m + 2) * (m + 2) * (m + 2), M3 = (long)m * m * m;
        if (l == H->nlev - 1) {
            h_coarse_cg(H, W, l, inv_dt);
            break;
        }
        memset(W[l].Z, 0, sizeof(double) * (size_t)nb * P);
        for (int s = 0; s < 2; s++) {
            h_exchange(L, W[l].Z);
            memcpy(W[l].Z2, W[l].Z, sizeof(double) * (size_t)nb * P);
            h_apply(L, nb, inv_dt, W[l].Z, W[l].R, W[l].Z2, 2, 0.8);
            double *t = W[l].Z; W[l].Z = W[l].Z2; W[l].Z2 = t;
        }
        h_exchange(L, W[l].Z);
        h_apply(L, nb, inv_dt, W[l].Z, W[l].R, W[l].RES, 1, 0.0);
        const int mc = m / 2;
        const long Mc = (long)mc * mc * mc;
        for (int b = 0; b < nb; b++)
            for (int k = 0; k < mc; k++) for (int j = 0; j < mc; j++) for (int i = 0; i < mc; i++) {
                double t = 0.0;
                for (int di = 0; di < 2; di++) for (int dj = 0; dj < 2; dj++) for (int dk = 0; dk < 2; dk++)
                    t += W[l].RES[(long)b * M3 + (2 * i + di) + (long)m * ((2 * j + dj) + (long)m * (2 * k + dk))];
                W[l + 1].R[(long)b * Mc + i + (long)mc * (j + (long)mc * k)] = t;
            }
    }
    for (int l = H->nlev - 2; l >= 0; l--) {
        const hlev *L = &H->L[l];
        const int m = L->m;
        const long P = (long)(m + 2) * (m + 2) * (m + 2);
        h_exchange(&H->L[l + 1], W[l + 1].Z);
        h_prolong_linear(H, l, W[l + 1].Z, W[l].Z);
        for (int s = 0; s < 2; s++) {
            h_exchange(L, W[l].Z);
            memcpy(W[l].Z2, W[l].Z, sizeof(double) * (size_t)nb * P);
            h_apply(L, nb, inv_dt, W[l].Z, W[l].R, W[l].Z2, 2, 0.8);
            double *t = W[l].Z; W[l].Z = W[l].Z2; W[l].Z2 = t;
        }
    }
    const int m = H->L[0].m;
    memcpy(z0_out, W[0].Z, sizeof(double) * (size_t)nb * (m + 2) * (m + 2) * (m + 2));
}

/* zf += P zc between level l+1 and level l (interior arrays, Fortran order
 * per block): the V-cycle's prolongation with the coarse level's ghost fill */
void orc_heat_prolong(orc_heat *H, int l, const double *zc, double *zf) {
    const int nb = H->nb, m = H->L[l].m, mc = m / 2;
    const long P = (long)(m + 2) * (m + 2) * (m + 2), Pc = (long)(mc + 2) * (mc + 2) * (mc + 2);
    const long M3 = (long)m * m * m, Mc = (long)mc * mc * mc;
    double *cp = malloc(sizeof(double) * (size_t)nb * Pc), *fp = malloc(sizeof(double) * (size_t)nb * P);
    h_pad(mc, nb, zc, cp);
    h_pad(m, nb, zf, fp);
    h_exchange(&H->L[l + 1], cp);
    h_prolong_linear(H, l, cp, fp);
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < m; k++) for (int j = 0; j < m; j++) for (int i = 0; i < m; i++)
            zf[(long)b * M3 + i + (long)m * (j + (long)m * k)] = fp[(long)b * P + HP(m, i + 1, j + 1, k + 1)];
    (void)Mc;
    free(cp); free(fp);
}

void orc_heat_set_coarse(orc_heat *H, double tol, int max_iter) {
    H->coarse_tol = tol;
    H->coarse_max = max_iter;
}

/* pad/unpad helpers (interior Fortran order per block) */
static void h_pad(int m, int nb, const double *in, double *out) {
    const long M3 = (long)m * m * m, P = (long)(m + 2) * (m + 2) * (m + 2);
    memset(out, 0, sizeof(double) * (size_t)nb * P);
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < m; k++) for (int j = 0; j < m; j++) for (int i = 0; i < m; i++)
            out[(long)b * P + HP(m, i + 1, j + 1, k + 1)] = in[(long)b * M3 + i + (long)m * (j + (long)m * k)];
}

void orc_heat_apply(orc_heat *H, double dt, const double *x, double *y) {
    const hlev *L = &H->L[0];
    const int m = L->m;
    double *xp = malloc(sizeof(double) * (size_t)H->nb * (m + 2) * (m + 2) * (m + 2));
    h_pad(m, H->nb, x, xp);
    h_exchange(L, xp);
    h_apply(L, H->nb, 1.0 / dt, xp, NULL, y, 0, 0.0);
    free(xp);
}

/* B = -grad(phi) at every interior cell, the gradient of
 * ModSpherical_Grad_f_O2 (3_6/ModSpherical.f90:161-199, second-order centred
 * differences (1./2./dx)*(f(+1) - f(-1)), then 1/r and 1/sin(theta)) with the
 * minus sign that 3_5/ModPFSS.f90:141 states (B = -grad Phi): with the flux
 * right-hand side b = Br S of the bottom faces, -dPhi/dr = Br there.  Face
 * ghosts as the operator's; at the domain's radial boundaries, the ghosts that
 * the boundary conditions give: bottom Phi_0 = Phi_1 + Br dr, top (Phi = 0 on
 * the face) Phi_{m+1} = -Phi_m.  br: for each bottom block in block order,
 * Br at its bottom faces (j, k), j fastest.  B: nblocks x 3 x (m,m,m). */
void orc_pfss_field(orc_heat *H, const double *phi, const double *br, double *B) {
    const hlev *L = &H->L[0];
    const int m = L->m, nb = H->nb;
    const long M3 = (long)m * m * m, P = (long)(m + 2) * (m + 2) * (m + 2);
    double *xp = malloc(sizeof(double) * (size_t)nb * P);
    h_pad(m, nb, phi, xp);
    h_exchange(L, xp);
    double *xc[3], *xf[3];
    for (int a = 0; a < 3; a++) { xc[a] = malloc(sizeof(double) * (m + 2)); xf[a] = malloc(sizeof(double) * (m + 1)); }
    int nbot = 0;
    for (int ib = 0; ib < nb; ib++) {
        const oblock *b = &H->c->b[ib];
        double *X = xp + (long)ib * P;
        lev_centres(b, m, xc, xf);
        double dx[3];
        for (int a = 0; a < 3; a++) dx[a] = (b->rng[a][1] - b->rng[a][0]) / (double)m;
        if (b->bottom) {
            for (int k = 1; k <= m; k++) for (int j = 1; j <= m; j++)
                X[HP(m, 0, j, k)] = X[HP(m, 1, j, k)] + br[(long)nbot * m * m + (long)(k - 1) * m + (j - 1)] * dx[0];
            nbot++;
        }
        if (b->top)
            for (int k = 1; k <= m; k++) for (int j = 1; j <= m; j++) X[HP(m, m + 1, j, k)] = -X[HP(m, m, j, k)];
        const double cr = (double)(1.0f / 2.0f) / dx[0], ct = (double)(1.0f / 2.0f) / dx[1], cp = (double)(1.0f / 2.0f) / dx[2];
        double *Br = B + (long)ib * 3 * M3, *Bt = Br + M3, *Bp = Bt + M3;
        for (int k = 1; k <= m; k++) for (int j = 1; j <= m; j++) for (int i = 1; i <= m; i++) {
            const long q = (long)(i - 1) + (long)m * ((long)(j - 1) + (long)m * (k - 1));
            const double rinv = 1.0 / xc[0][i], sinv = 1.0 / sin(xc[1][j]);
            const double gr = cr * (X[HP(m, i + 1, j, k)] - X[HP(m, i - 1, j, k)]);
            double gt = ct * (X[HP(m, i, j + 1, k)] - X[HP(m, i, j - 1, k)]);
            gt = gt * rinv;
            double gp = cp * (X[HP(m, i, j, k + 1)] - X[HP(m, i, j, k - 1)]);
            gp = gp * rinv;
            gp = gp * sinv;
            Br[q] = -gr; Bt[q] = -gt; Bp[q] = -gp;
        }
    }
    for (int a = 0; a < 3; a++) { free(xc[a]); free(xf[a]); }
    free(xp);
}

static void h_work_alloc(orc_heat *H, hwork *W) {
    for (int l = 0; l < H->nlev; l++) {
        int ml = H->L[l].m;
        long Pl = (long)(ml + 2) * (ml + 2) * (ml + 2), Ml = (long)ml * ml * ml;
        W[l].Z = calloc((size_t)H->nb * Pl, sizeof(double)); W[l].Z2 = calloc((size_t)H->nb * Pl, sizeof(double));
        W[l].R = l ? calloc((size_t)H->nb * Ml, sizeof(double)) : NULL; W[l].RES = calloc((size_t)H->nb * Ml, sizeof(double));
    }
}

static void h_work_free(orc_heat *H, hwork *W) {
    W[0].R = NULL;  /* level 0 aliases the caller's residual during the V-cycle */
    for (int l = 0; l < H->nlev; l++) { free(W[l].Z); free(W[l].Z2); free(W[l].R); free(W[l].RES); }
}

void orc_heat_vcycle(orc_heat *H, double dt, const double *r, double *z) {
    const int nb = H->nb, m = H->L[0].m;
    const long M3 = (long)m * m * m, P = (long)(m + 2) * (m + 2) * (m + 2);
    hwork W[8];
    h_work_alloc(H, W);
    double *zp = malloc(sizeof(double) * (size_t)nb * P);
    h_vcycle(H, W, 1.0 / dt, r, zp);
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < m; k++) for (int j = 0; j < m; j++) for (int i = 0; i < m; i++)
            z[(long)b * M3 + i + (long)m * (j + (long)m * k)] = zp[(long)b * P + HP(m, i + 1, j + 1, k + 1)];
    free(zp);
    h_work_free(H, W);
}

void orc_heat_diag(orc_heat *H, double dt, double *d) {
    const hlev *L = &H->L[0];
    const long N = (long)H->nb * L->m * L->m * L->m;
    const double inv_dt = 1.0 / dt;
    for (long c = 0; c < N; c++) d[c] = L->V[c] * inv_dt + L->off[c];
}

static double h_dot(int m, int nb, const double *a, int apad, const double *b, int bpad) {
    const long M3 = (long)m * m * m, P = (long)(m + 2) * (m + 2) * (m + 2);
    double s = 0.0;
    for (int blk = 0; blk < nb; blk++)
        for (int k = 0; k < m; k++) for (int j = 0; j < m; j++) for (int i = 0; i < m; i++) {
            long ii = (long)blk * M3 + i + (long)m * (j + (long)m * k);
            long pp = (long)blk * P + HP(m, i + 1, j + 1, k + 1);
            s += (apad ? a[pp] : a[ii]) * (bpad ? b[pp] : b[ii]);
        }
    return s;
}

/* Solve A T = (V/dt) T^n to ||b - A T||/||b|| <= tol by FCG(1) with the V-cycle.
 * Returns iterations; x_out interior (Fortran order per block). */
int orc_heat_solve(orc_heat *H, double dt, const double *Tn, double tol, int max_iter, double *x_out, double *relres) {
    const int nb = H->nb, m = H->L[0].m;
    const long N = (long)nb * m * m * m;
    const double inv_dt = 1.0 / dt;
    double *b = malloc(sizeof(double) * N);
    for (long c = 0; c < N; c++) b[c] = H->L[0].V[c] * inv_dt * Tn[c];
    int it = orc_heat_solve_b(H, dt, b, tol, max_iter, x_out, relres);
    free(b);
    return it;
}

/* A x = b0 from x = 0 by the same flexible CG */
int orc_heat_solve_b(orc_heat *H, double dt, const double *b0, double tol, int max_iter, double *x_out, double *relres) {
    const int nb = H->nb, m = H->L[0].m;
    const long M3 = (long)m * m * m, P = (long)(m + 2) * (m + 2) * (m + 2), N = (long)nb * M3;
    const double inv_dt = 1.0 / dt;
    hwork W[8];
    h_work_alloc(H, W);
    double *b = malloc(sizeof(double) * N), *r = malloc(sizeof(double) * N), *x = calloc(N, sizeof(double));
    double *q = malloc(sizeof(double) * N), *p = calloc((size_t)nb * P, sizeof(double)), *z = malloc(sizeof(double) * nb * P);
    for (long c = 0; c < N; c++) { b[c] = b0[c]; r[c] = b[c]; }
    double bn = h_dot(m, nb, b, 0, b, 0);
    if (bn == 0.0) {  /* b = 0: the solution is 0 */
        memset(x_out, 0, sizeof(double) * N);
        if (relres) *relres = 0.0;
        h_work_free(H, W);
        free(b); free(r); free(x); free(q); free(p); free(z);
        return 0;
    }
    h_vcycle(H, W, inv_dt, r, z);
    for (long c = 0; c < (long)nb * P; c++) p[c] = z[c];
    int it = 0;
    double rel = 1.0, gamma = 1.0;
    while (1) {
        h_exchange(&H->L[0], p);
        h_apply(&H->L[0], nb, inv_dt, p, NULL, q, 0, 0.0);
        gamma = h_dot(m, nb, p, 1, q, 0);
        double alpha = h_dot(m, nb, p, 1, r, 0) / gamma;
        for (int blk = 0; blk < nb; blk++)
            for (long c = 0; c < M3; c++) {
                int i = (int)(c % m), j = (int)((c / m) % m), k = (int)(c / ((long)m * m));
                x[blk * M3 + c] += alpha * p[(long)blk * P + HP(m, i + 1, j + 1, k + 1)];
                r[blk * M3 + c] -= alpha * q[blk * M3 + c];
            }
        it++;
        rel = sqrt(h_dot(m, nb, r, 0, r, 0) / bn);
        if (rel <= tol || it >= max_iter) break;
        h_vcycle(H, W, inv_dt, r, z);
        double beta = h_dot(m, nb, z, 1, q, 0) / gamma;
        for (long c = 0; c < (long)nb * P; c++) p[c] = z[c] - beta * p[c];
    }
    memcpy(x_out, x, sizeof(double) * N);
    if (relres) *relres = rel;
    h_work_free(H, W);
    free(b); free(r); free(x); free(q); free(p); free(z);
    return it;
}
