// The learned surrogate: a 1-D Fourier neural operator trained and run on the device.
//
// It maps a scan point (its input spectrum on n bins and its scan parameters) to the
// exit spectrum the solver would produce, so that a scan can be screened before the
// solver runs (the project's own component; no counterpart in the reference). Every
// dense product runs on the matrix cores in BF16 with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16): the 1x1 convolutions, the truncated DFT onto the first
// `modes` bins and its inverse (dense matrices, K = n and K = 2 modes), the head. The
// per-mode complex mixing, the loss, the reductions and Adam run in FP32.
//
// Layouts: activations (batch, channels, n) row-major; parameters one FP32 vector in the
// order of oracle/fno.py's `layout` (lift, then per layer the spectral weights' real and
// imaginary parts (modes, width_in, width_out), the convolution (width_out, width_in)
// and its bias, then the head), mirrored in BF16 for the products.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace sp::fno {

using bf16 = std::uint16_t;

struct Shape {
  int n_params = 4;  // scan parameters; the input channels are the spectrum, these, the grid
  int width = 64;
  int modes = 64;
  int layers = 4;
  int hidden = 128;
  int n = 2048;
  int in_ch() const { return n_params + 2; }
};

// Offsets of the parameter tensors in the flat vector.
struct Offsets {
  long lift_w, lift_b;
  std::vector<long> r_re, r_im, w, b;
  long q1_w, q1_b, q2_w, q2_b;
  long total;
};
Offsets offsets(const Shape& s);

class Network {
 public:
  Network(const Shape& s, int max_batch);
  ~Network();
  Network(const Network&) = delete;
  Network& operator=(const Network&) = delete;

  long parameter_count() const { return off_.total; }
  void set_parameters(const float* host);
  void get_parameters(float* host);
  // Outputs (batch, n) of host inputs: spectra (batch, n), params (batch, n_params).
  void forward(const float* spectra, const float* params, int batch, float* out);
  // The mean squared error against `target` (batch, n) and its gradient (host, flat).
  double gradient(const float* spectra, const float* params, const float* target, int batch,
                  float* grad);
  // A device-resident training set, then Adam steps over minibatches of its rows
  // (`indices`, steps x batch, drawn by the caller); each step's loss into `losses`.
  void load_dataset(const float* spectra, const float* params, const float* targets, long n);
  void train(const int* indices, long steps, int batch, float lr, float* losses);
  void reset_optimizer();
  // The last train call's device time, ms.
  double train_ms() const { return train_ms_; }

 private:
  void upload_batch(const float* spectra, const float* params, const float* target, int batch);
  void run_forward(int batch);
  void run_backward(int batch);  // needs d_out_ set; fills grad_
  void zero_grad();
  Shape s_;
  Offsets off_;
  int max_batch_;
  hipStream_t stream_ = nullptr;
  std::vector<void*> owned_;
  template <typename T>
  T* alloc(std::size_t n);
  // Parameters, their BF16 mirror, gradient and Adam moments.
  float *param_ = nullptr, *grad_ = nullptr, *m_ = nullptr, *v_ = nullptr;
  bf16* param16_ = nullptr;
  long adam_t_ = 0;
  // The truncated transforms (BF16): F (n, 2 modes), G (2 modes, n).
  bf16 *f16_ = nullptr, *g16_ = nullptr;
  // The batch: inputs, target, activations and their gradients.
  float *spec_ = nullptr, *prm_ = nullptr, *tgt_ = nullptr, *out_ = nullptr;
  std::vector<bf16*> v16_;       // layers + 1, (B, W, n)
  std::vector<float*> z_;        // layers, (B, W, n) pre-activation
  std::vector<float*> vhat_;     // layers, (B, W, 2 modes)
  bf16* yhat16_ = nullptr;       // (B, W, 2 modes)
  float *h1_ = nullptr;          // (B, H, n)
  bf16* h16_ = nullptr;
  float *dout_ = nullptr, *dv_ = nullptr, *dz_ = nullptr, *dyhat_ = nullptr, *dh1_ = nullptr;
  bf16 *dz16_ = nullptr, *dvhat16_ = nullptr, *dh116_ = nullptr, *dout16_ = nullptr;
  float* partials_ = nullptr;    // loss partial sums
  // Split-K shares of the long-K products (weight gradients, the truncated transforms).
  static constexpr long kSplitFloats = 16L << 20;
  float* split_ = nullptr;
  float* losses_ = nullptr;      // per training step
  long losses_cap_ = 0;
  int* indices_ = nullptr;
  long indices_cap_ = 0;
  // Dataset.
  float *ds_spec_ = nullptr, *ds_prm_ = nullptr, *ds_tgt_ = nullptr;
  long ds_n_ = 0;
  double train_ms_ = 0.0;
};

// C = A . B on the matrix cores (testing): host FP32 operands rounded to BF16, FP32 out.
void testing_gemm_bf16(int M, int N, int K, const float* A, const float* B, float* C);

}  // namespace sp::fno
