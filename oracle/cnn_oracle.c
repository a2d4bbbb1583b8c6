/* CNN oracle: forward and backward of the 1->16->16->1 3^3 network, FP64.
 * TEST INFRASTRUCTURE ONLY.  Conventions in cnn_oracle.h. */
#include "cnn_oracle.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

float cnn_round_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u = u + 0x7fffu + ((u >> 16) & 1u);   /* round to nearest, ties to even */
    u &= 0xffff0000u;
    float y;
    memcpy(&y, &u, 4);
    return y;
}

static double rb(double x, int bf16) { return bf16 ? (double)cnn_round_bf16((float)x) : x; }

/* neighbour of voxel v through tap t, or -1 outside the block */
static int nbr(int v, int t) {
    int i = v >> 8, j = (v >> 4) & 15, k = v & 15;
    i += t / 9 - 1; j += (t / 3) % 3 - 1; k += t % 3 - 1;
    if (i < 0 || i > 15 || j < 0 || j > 15 || k < 0 || k > 15) return -1;
    return (i << 8) | (j << 4) | k;
}

static void weights(const float *p, int bf16, double *w) {
    for (int q = 0; q < CNN_NPARAM; q++) {
        int is_bias = (q >= CNN_OFF_B1 && q < CNN_OFF_W2) || (q >= CNN_OFF_B2 && q < CNN_OFF_W3) || q == CNN_OFF_B3;
        w[q] = is_bias ? (double)p[q] : rb(p[q], bf16);
    }
}

void cnn_forward(const float *params, int nimg, const float *x, double *h1o, double *h2o, double *out, int bf16) {
    double *w = malloc(sizeof(double) * CNN_NPARAM);
    weights(params, bf16, w);
    const double *W1 = w + CNN_OFF_W1, *B1 = w + CNN_OFF_B1, *W2 = w + CNN_OFF_W2, *B2 = w + CNN_OFF_B2,
                 *W3 = w + CNN_OFF_W3, *B3 = w + CNN_OFF_B3;
#pragma omp parallel for schedule(static) if (nimg > 1)
    for (int n = 0; n < nimg; n++) {
        double *xl = malloc(sizeof(double) * CNN_V), *a = malloc(sizeof(double) * CNN_V * CNN_C),
               *b = malloc(sizeof(double) * CNN_V * CNN_C);
        for (int v = 0; v < CNN_V; v++) xl[v] = rb(x[(long)n * CNN_V + v], bf16);
        for (int v = 0; v < CNN_V; v++)
            for (int co = 0; co < CNN_C; co++) {
                double s = 0.0;
                for (int t = 0; t < 27; t++) {
                    int u = nbr(v, t);
                    if (u >= 0) s += W1[t * 16 + co] * xl[u];
                }
                s += B1[co];
                a[v * 16 + co] = rb(s > 0.0 ? s : 0.0, bf16);
            }
        for (int v = 0; v < CNN_V; v++)
            for (int co = 0; co < CNN_C; co++) {
                double s = 0.0;
                for (int t = 0; t < 27; t++) {
                    int u = nbr(v, t);
                    if (u < 0) continue;
                    for (int ci = 0; ci < 16; ci++) s += W2[(t * 16 + ci) * 16 + co] * a[u * 16 + ci];
                }
                s += B2[co];
                b[v * 16 + co] = rb(s > 0.0 ? s : 0.0, bf16);
            }
        for (int v = 0; v < CNN_V; v++) {
            double s = 0.0;
            for (int t = 0; t < 27; t++) {
                int u = nbr(v, t);
                if (u < 0) continue;
                for (int ci = 0; ci < 16; ci++) s += W3[t * 16 + ci] * b[u * 16 + ci];
            }
            out[(long)n * CNN_V + v] = s + B3[0];
        }
        if (h1o) memcpy(h1o + (long)n * CNN_V * CNN_C, a, sizeof(double) * CNN_V * CNN_C);
        if (h2o) memcpy(h2o + (long)n * CNN_V * CNN_C, b, sizeof(double) * CNN_V * CNN_C);
        free(xl); free(a); free(b);
    }
    free(w);
}

void cnn_backward(const float *params, int nimg, const float *x, const double *h1, const double *h2,
                  const double *gout, double *grad, int bf16) {
    double *w = malloc(sizeof(double) * CNN_NPARAM);
    weights(params, bf16, w);
    const double *W2 = w + CNN_OFF_W2, *W3 = w + CNN_OFF_W3;
    memset(grad, 0, sizeof(double) * CNN_NPARAM);
    double *g3 = malloc(sizeof(double) * CNN_V), *g2 = malloc(sizeof(double) * CNN_V * CNN_C),
           *g1 = malloc(sizeof(double) * CNN_V * CNN_C), *xl = malloc(sizeof(double) * CNN_V);
    for (int n = 0; n < nimg; n++) {
        const double *H1 = h1 + (long)n * CNN_V * CNN_C, *H2 = h2 + (long)n * CNN_V * CNN_C;
        for (int v = 0; v < CNN_V; v++) {
            g3[v] = rb(gout[(long)n * CNN_V + v], bf16);
            xl[v] = rb(x[(long)n * CNN_V + v], bf16);
            grad[CNN_OFF_B3] += gout[(long)n * CNN_V + v];
        }
        /* layer 3: dW3[t][ci] = sum_v h2[v+t][ci] g3[v]; g2[u][ci] = sum over (v,t) with v+t = u */
        memset(g2, 0, sizeof(double) * CNN_V * CNN_C);
        for (int v = 0; v < CNN_V; v++)
            for (int t = 0; t < 27; t++) {
                int u = nbr(v, t);
                if (u < 0) continue;
                for (int ci = 0; ci < 16; ci++) {
                    grad[CNN_OFF_W3 + t * 16 + ci] += H2[u * 16 + ci] * g3[v];
                    g2[u * 16 + ci] += W3[t * 16 + ci] * g3[v];
                }
            }
        for (int q = 0; q < CNN_V * CNN_C; q++) g2[q] = rb(H2[q] > 0.0 ? g2[q] : 0.0, bf16);
        for (int q = 0; q < CNN_V * CNN_C; q++) grad[CNN_OFF_B2 + (q & 15)] += g2[q];
        /* layer 2 */
        memset(g1, 0, sizeof(double) * CNN_V * CNN_C);
        for (int v = 0; v < CNN_V; v++)
            for (int t = 0; t < 27; t++) {
                int u = nbr(v, t);
                if (u < 0) continue;
                for (int ci = 0; ci < 16; ci++)
                    for (int co = 0; co < 16; co++) {
                        grad[CNN_OFF_W2 + (t * 16 + ci) * 16 + co] += H1[u * 16 + ci] * g2[v * 16 + co];
                        g1[u * 16 + ci] += W2[(t * 16 + ci) * 16 + co] * g2[v * 16 + co];
                    }
            }
        for (int q = 0; q < CNN_V * CNN_C; q++) g1[q] = rb(H1[q] > 0.0 ? g1[q] : 0.0, bf16);
        for (int q = 0; q < CNN_V * CNN_C; q++) grad[CNN_OFF_B1 + (q & 15)] += g1[q];
        /* layer 1 */
        for (int v = 0; v < CNN_V; v++)
            for (int t = 0; t < 27; t++) {
                int u = nbr(v, t);
                if (u < 0) continue;
                for (int co = 0; co < 16; co++) grad[CNN_OFF_W1 + t * 16 + co] += xl[u] * g1[v * 16 + co];
            }
    }
    free(g3); free(g2); free(g1); free(xl); free(w);
}
