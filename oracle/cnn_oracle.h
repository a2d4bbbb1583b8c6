/* CNN oracle: the learned V-cycle correction's network on the CPU, FP64.
 * TEST INFRASTRUCTURE ONLY (tests/ and the checkers; never the product).
 *
 * The reference has no learned component.  The network is defined here:
 * three 3x3x3 convolutions 1 -> 16 -> 16 -> 1 over one 16^3 block image with
 * zero padding at the block's faces, ReLU after the first two, biases.
 * Tap t = (di+1)*9 + (dj+1)*3 + (dk+1) reads voxel (i+di, j+dj, k+dk);
 * voxel index v = (i*16 + j)*16 + k.
 *
 * Parameters, one flat float array of CNN_NPARAM (= 7809):
 *   W1[27][16] (t, co), b1[16], W2[27][16][16] (t, ci, co), b2[16],
 *   W3[27][16] (t, ci), b3[1].
 * Activations: x[img][4096] (float), h1/h2[img][4096][16] (post-ReLU),
 * out[img][4096].
 *
 * bf16 != 0 reproduces the GPU's rounding points: x, weights, h1, h2 and the
 * backward's MFMA operands (g3, g2, g1) are rounded to bfloat16
 * (round-to-nearest-even); biases and accumulations stay wide.  bf16 == 0 is
 * the plain FP64 network (used for finite-difference gradient checks). */
#ifndef FALCON_CNN_ORACLE_H
#define FALCON_CNN_ORACLE_H

#define CNN_E 16
#define CNN_V 4096
#define CNN_C 16
#define CNN_NPARAM 7809
#define CNN_OFF_W1 0
#define CNN_OFF_B1 432
#define CNN_OFF_W2 448
#define CNN_OFF_B2 7360
#define CNN_OFF_W3 7376
#define CNN_OFF_B3 7808

float cnn_round_bf16(float x);
/* h1, h2 may be NULL when not wanted */
void cnn_forward(const float *params, int nimg, const float *x, double *h1, double *h2, double *out, int bf16);
/* gradient of L = sum_{img,v} gout[img][v] * out[img][v] with respect to the
 * parameters (grad, CNN_NPARAM doubles) given the forward activations */
void cnn_backward(const float *params, int nimg, const float *x, const double *h1, const double *h2,
                  const double *gout, double *grad, int bf16);
#endif
