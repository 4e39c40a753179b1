// Warm-start network (mlp.hip).
#pragma once
#include "common.hpp"

namespace cn {

class Mlp {
  public:
    static constexpr int kInPad = 16;  // inputs are zero-padded to one MFMA K step

    Mlp(int n_in, int n_hidden, int n_out, unsigned seed, hipStream_t s);
    long n_params() const { return n_params_; }
    int n_in() const { return n_in_; }
    int n_out() const { return n_out_; }
    // Flat FP32 parameters: W1 (16 x H, rows >= n_in zero), b1, W2 (H x H), b2, W3 (H x O), b3.
    void set_params(const float* host);
    void get_params(float* host) const;
    // Mean squared error of one batch and the gradients of every parameter (same layout).
    double gradients(const float* X, const float* T, int B, float* grads_host);
    // Adam over `epochs` passes of shuffled minibatches; returns the last epoch's mean loss.
    double train(const float* X, const float* T, long N, int epochs, int batch, float lr, unsigned seed);
    // BF16 MFMA inference: Y (B, n_out) FP32.
    void predict(const float* X, int B, float* Y_host);

  private:
    void ensure_batch(int B);
    void upload_inputs(const float* X, int B);
    double forward_backward(int B);

    int n_in_, n_hidden_, n_out_;
    hipStream_t s_;
    long n_params_ = 0;
    long offset_[3] = {}, boffset_[3] = {};
    long steps_ = 0;
    int cap_batch_ = 0;
    DeviceBuffer<float> w_, g_, m_, v_;
    DeviceBuffer<unsigned short> w16_;
    DeviceBuffer<float> x_, z1_, h1_, z2_, h2_, y_, t_, dy_, dh_, dz_;
    DeviceBuffer<double> loss_;
    PinnedStage stage_;  // padded inputs, copied without a wait
    DeviceBuffer<unsigned short> x16_, h16a_, h16b_;
};

}  // namespace cn
