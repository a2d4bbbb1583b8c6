// Learned V-cycle correction: a 1 -> 16 -> 16 -> 1 network of 3x3x3
// convolutions over one 16^3 block image (zero padding at the block faces,
// ReLU after the first two layers), BF16 weights and activations on MFMA
// (implicit GEMM, FP32 accumulate), FP32 master weights trained with Adam by
// hand-written forward and backward kernels (cnn.hip).
//
// Parameter vector (7809 floats): W1[27][16] (t, co), b1[16],
// W2[27][16][16] (t, ci, co), b2[16], W3[27][16] (t, ci), b3[1];
// tap t = (di+1)*9 + (dj+1)*3 + (dk+1), voxel v = (i*16 + j)*16 + k.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "heat.hpp"

namespace falcon {

constexpr int kCnnParams = 7809;

class CnnModel : public HeatCorrector {
  public:
    CnnModel();
    ~CnnModel() override;
    void init(uint64_t seed);                        // He-normal weights, zero biases, Adam state cleared
    void set_params(const float *p);                 // host, kCnnParams
    void get_params(float *p) const;
    // out[img][4096] = net(x[img][4096]) (device pointers, stream-ordered)
    void infer(const float *x, float *out, long long nimg, hipStream_t stream) override;
    void infer_heat(const float *x, const double *scale, int scale_stride, int nb, long long istride, double *c,
                    long long nimg, const int *active, hipStream_t stream) override;
    void prepare(hipStream_t stream) override { pack(stream); }
    // sum over the batch of the gradient of sum(gout * out) (device x/gout, host grad)
    void gradient(const float *x, const float *gout, int nimg, double *grad_host, hipStream_t stream);
    // one Adam step on the mean squared error over a batch (device x, y); returns the loss
    double train_step(const float *x, const float *y, int nimg, double lr, hipStream_t stream);
    // epochs over nimg device images in shuffled minibatches, learning rate decaying
    // linearly to lr/10; returns the last epoch's mean loss
    double train(const float *x, const float *y, long long nimg, int epochs, int batch, double lr, uint64_t seed,
                 hipStream_t stream);
    int max_batch() const { return max_batch_; }
    // training loss: alpha < 0 the mean squared error (default); alpha >= 0 the
    // gradient-weighted error alpha |d|^2 + sum_a w_a |d_a d|^2 per image, the
    // face weights w per image axis (r, theta, phi) (k_cnn_loss_h1)
    // scale_free: the per-image fraction of e's norm that the best multiple of
    // the output leaves (k_cnn_loss_cos) instead of the norm of the difference
    void set_loss_h1(double alpha, double wr, double wt, double wp, bool scale_free) {
        h1_alpha_ = alpha;
        h1_w_[0] = wr; h1_w_[1] = wt; h1_w_[2] = wp;
        loss_cos_ = scale_free;
    }

  private:
    void ensure_batch(int nimg);
    void pack(hipStream_t stream);
    void backward(const float *x, int nimg, hipStream_t stream);   // from g3 in ws_g3_ into grad_
    float *params_ = nullptr, *adam_m_ = nullptr, *adam_v_ = nullptr;
    double *grad_ = nullptr;
    void *packed_ = nullptr;      // MFMA operand images of the weights (fwd and data-gradient forms)
    bool packed_valid_ = false;
    int max_batch_ = 0;
    uint16_t *h1_ = nullptr, *h2_ = nullptr, *g2_ = nullptr, *g1_ = nullptr;
    float *out_ = nullptr, *g3_ = nullptr, *part_ = nullptr, *xb_ = nullptr, *yb_ = nullptr;
    double *loss_part_ = nullptr;
    long long adam_t_ = 0;
    double h1_alpha_ = -1.0, h1_w_[3] = {1.0, 1.0, 1.0};
    bool loss_cos_ = false;
};

}  // namespace falcon
