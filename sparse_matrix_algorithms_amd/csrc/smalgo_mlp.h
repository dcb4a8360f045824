// FP32 two-layer perceptron with hand-written HIP forward/backward/Adam.
#pragma once

#include "smalgo_solver.h"

namespace smalgo {

struct Mlp {
  Ctx& cx;
  int nin, nh, nout;
  size_t np = 0;          // parameter count: W1, b1, W2, b2 (in that order)
  DBuf<float> P, G, M, Vv;  // parameters, gradient, Adam moments
  DBuf<float> X, T, H, Y, dY, dH, L;
  int cap = 0;
  long long step = 0;
  Mlp(Ctx& c, int nin, int nh, int nout, uint64_t seed);
  void set_params(const float* host);
  void get_params(float* host) const;
  void predict(const float* X, int B, float* Y);
  // loss = mean((Y-T)^2) over the batch; g (optional) receives dloss/dparams
  float gradient(const float* X, const float* T, int B, float* g);
  // full-batch Adam; returns the last loss (loss_hist, if given, gets every epoch's)
  float train(const float* X, const float* T, int B, int epochs, float lr, float* loss_hist);

 private:
  void ensure_batch(int B);
  void load_batch(const float* X, const float* T, int B);
  void forward_dev(int B);
  float backward_dev(int B);
};

}  // namespace smalgo
