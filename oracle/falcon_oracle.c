/* Oracle: CPU FP64 restatement of the Falcon reference RK4 hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see falcon_oracle.h).  Each function cites the
 * reference routine it restates; operation order and the single-precision
 * literals of the reference (3_6/ModConst.f90, `1./12.`, `dpi*0.25`, ...) are
 * reproduced so results can be compared with the flang build of the reference.
 * Compiled with -ffp-contract=off: no fused multiply-add anywhere.
 */
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
    double *u, *urk;        /* padded x 8 */
    double *r;              /* interior x 8 */
} oblock;

struct orc_ctx {
    int ni, nj, nk, ng;
    int NI, NJ, NK;          /* padded extents */
    double r_range[2];
    onode yin, yang;
    int nleaf, nleaf_yin;
    oblock *b;
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

/* ModYinYang_CoordConv_0D (3_6/ModYinYang.f90:8-17) */
static void coordconv(const double in[3], double out[3]) {
    double st = sin(in[1]), sp = sin(in[2]), cp = cos(in[2]), ct = cos(in[1]);
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
        b->u = calloc(np * ORC_NVAR, sizeof(double));
        b->urk = calloc(np * ORC_NVAR, sizeof(double));
        b->r = calloc(nint * ORC_NVAR, sizeof(double));
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
        free(b->u); free(b->urk); free(b->r);
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
static void interp_point(const orc_ctx *c, const oblock *s, const double *f, const double x[3], double out[ORC_NVAR]) {
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
    for (int v = 0; v < ORC_NVAR; v++) {
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

/* ModYinYang_VecConv_1D (3_6/ModYinYang.f90:48-61) for one point */
static void vecconv(const double rtp[3], double v[3]) {
    double o[3];
    coordconv(rtp, o);
    double sp = sin(rtp[2]), spo = sin(o[2]), cp = cos(rtp[2]), sto = sin(o[1]);
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
                double val[ORC_NVAR];
                interp_point(c, s, f, x, val);
                if (sc->recv_yin != s->yin) {
                    double bv[3] = {val[4], val[5], val[6]}, vv[3] = {val[1], val[2], val[3]};
                    vecconv(x, bv);
                    vecconv(x, vv);
                    val[4] = bv[0]; val[5] = bv[1]; val[6] = bv[2];
                    val[1] = vv[0]; val[2] = vv[1]; val[3] = vv[2];
                }
                size_t q = PAD(c, sc->ijk[g], sc->ijk[g + sc->ngc], sc->ijk[g + 2 * sc->ngc]);
                for (int v = 0; v < ORC_NVAR; v++) ft[q + nv * v] = val[v];
            }
        }
    }
}

/* ------------------------------------------------------------------ physics */

enum { RHO1 = 0, VR = 1, VT = 2, VP = 3, BR = 4, BT = 5, BP = 6, S1 = 7 };

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
        double cot = cos(xj) / sin(xj);
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

/* ModDiffusion_Aritificial_1 (3_6/ModDiffusion.f90:9-150), h integer, all 8 vars;
 * MAX(0.0,NaN) pinned to 0.0. R accumulates (caller zeroes it). */
static void diffusion(const orc_ctx *c, const oblock *b, const double *p, int h, double *R) {
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
        for (int v = 0; v < 8; v++) {
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
    diffusion(c, b, b->u, 2, r_out);
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
            double x = b->h[INT(c, i, j, k)] / w;
            if (first || x < mn) { mn = x; first = 0; }
        }
        double dt1 = cfl * mn;
        dt = dt < dt1 ? dt : dt1;
    }
    return dt;
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
            boundary(c, b, in);
            equation(c, b, in, b->r);
            double div = (double)(5.0f - (float)rk);
            double *out = rk < 4 ? b->urk : b->u;
            for (int v = 0; v < 8; v++)
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
        memset(b->r, 0, sizeof(double) * nn * 8);
        diffusion(c, b, b->u, 2, b->r);
        for (int v = 0; v < 8; v++)
            for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
                size_t q = PAD(c, i, j, k) + np * v;
                b->u[q] = b->u[q] + (0.5 * dt) * b->r[INT(c, i, j, k) + nn * v];
            }
    }
    orc_exchange(c, 0);
    return dt;
}

/* Yin-Yang point map and the VecConv coefficients (sin p sin p', cos p / sin t')
 * for one donor-frame point, as used by vecconv() above. */
void orc_coordconv(const double in[3], double out[3]) { coordconv(in, out); }
void orc_vecconv_coeffs(const double x[3], double out[2]) {
    double o[3];
    coordconv(x, o);
    double sp = sin(x[2]), spo = sin(o[2]), cp = cos(x[2]), sto = sin(o[1]);
    out[0] = sp * spo;
    out[1] = cp / sto;
}
