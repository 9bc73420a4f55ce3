// CPU oracle of the learned WENO5 weights (an extension of the reference's
// FiniteDifferences::weno5, lsFiniteDifferences.hpp:88-123).
//
// TEST INFRASTRUCTURE ONLY (see ls_oracle.hpp): the parity checker of the
// GPU MLP inference inside the Hamiltonian kernel and of the GPU training
// kernels (loss and gradients).
//
// Network: inputs are the five one-sided differences d[0..4] of weno5,
// divided by max|d|; 5 -> 32 -> 32 -> 3, GELU (tanh approximation) after the hidden
// layers, softmax output = weights of weno5's three candidate polynomials.
// The weighted sum replaces sum(alpha_k p_k)/sum(alpha_k).  Where the
// regularised smoothness indicators (eps*dx^2 + beta_k) differ by more than a
// factor 10 the plain weno5 result is used (kink fallback).
//
// Parity status: mlpForward / mlpLossGrad are pinned by finite differences of
// the loss (tests/test_oracle_mlp.py) and by closed forms (zero network =>
// uniform weights; a network whose logits equal log of weno5's ideal
// weights reproduces the linear-weight scheme).  weno5L with the fallback
// forced is pinned against weno5.
#include "ls_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace lso {

float bf16Round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan: truncate (keep nan quiet)
    u &= 0xffff0000u;
    if ((u & 0x007f0000u) == 0 && (x != x)) u |= 0x00400000u;
  } else {
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;  // round to nearest even
    u &= 0xffff0000u;
  }
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

// GELU with the tanh approximation, 0.5 x (1 + tanh(u)), written as
// x sigmoid(2u), u = sqrt(2/pi) (x + 0.044715 x^3); derivative
// sigmoid(2u) + 2 x sigmoid(2u) (1 - sigmoid(2u)) u'.
template <class R> static inline R geluT(R x, R &dgelu) {
  const R c = (R)0.7978845608028654;  // sqrt(2/pi)
  const R a = (R)0.044715;
  const R x2 = x * x;
  const R u = c * (x + a * x2 * x);
  const R sg = (R)1 / ((R)1 + std::exp((R)-2 * u));
  dgelu = sg + (R)2 * x * sg * ((R)1 - sg) * c * ((R)1 + (R)3 * a * x2);
  return x * sg;
}

// Inference: P holds the network parameters as floats (the bf16 values the
// GPU uses); fp32 accumulation in index order.  The forward pass of the
// training kernel: layer 1 bias first, its activations rounded to bf16 for
// the 32x32 product, which is summed before its bias; layer 3 bias first.
void mlpForward(const float *P, const float in[MLP_IN], float w[MLP_OUT]) {
  float x[MLP_IN], h1[MLP_H], h2[MLP_H], z3[MLP_OUT];
  float dg;
  for (int i = 0; i < MLP_IN; ++i) x[i] = bf16Round(in[i]);
  for (int j = 0; j < MLP_H; ++j) {
    float z = P[MLP_B1 + j];
    for (int i = 0; i < MLP_IN; ++i) z += P[MLP_W1 + j * MLP_IN + i] * x[i];
    h1[j] = bf16Round(geluT<float>(z, dg));
  }
  for (int j = 0; j < MLP_H; ++j) {
    float z = 0.f;
    for (int i = 0; i < MLP_H; ++i) z += P[MLP_W2 + j * MLP_H + i] * h1[i];
    h2[j] = bf16Round(geluT<float>(z + P[MLP_B2 + j], dg));
  }
  for (int k = 0; k < MLP_OUT; ++k) {
    float z = P[MLP_B3 + k];
    for (int i = 0; i < MLP_H; ++i) z += P[MLP_W3 + k * MLP_H + i] * h2[i];
    z3[k] = z;
  }
  float m = std::max(z3[0], std::max(z3[1], z3[2]));
  float e[MLP_OUT], s = 0.f;
  for (int k = 0; k < MLP_OUT; ++k) {
    e[k] = std::exp(z3[k] - m);
    s += e[k];
  }
  for (int k = 0; k < MLP_OUT; ++k) w[k] = e[k] / s;
}

// Differences, smoothness indicators and candidate polynomials of weno5
// (lsFiniteDifferences.hpp:90-119).
struct Weno5Parts {
  double d[5];
  double s[3];
  double p[3];
};
static Weno5Parts weno5Parts(const double *x, bool plus) {
  Weno5Parts q;
  double *d = q.d;
  if (!plus) {
    for (int i = 0; i < 5; ++i) d[i] = x[i + 1] - x[i];
  } else {
    for (int i = 0; i < 5; ++i) d[i] = x[6 - i] - x[5 - i];
  }
  auto p2 = [](double v) { return v * v; };
  q.s[0] = (13.0 / 12.0) * p2(d[0] - 2 * d[1] + d[2]) + (1.0 / 4.0) * p2(d[0] - 4 * d[1] + 3 * d[2]);
  q.s[1] = (13.0 / 12.0) * p2(d[1] - 2 * d[2] + d[3]) + (1.0 / 4.0) * p2(d[1] - d[3]);
  q.s[2] = (13.0 / 12.0) * p2(d[2] - 2 * d[3] + d[4]) + (1.0 / 4.0) * p2(3 * d[2] - 4 * d[3] + d[4]);
  q.p[0] = (1.0 / 6.0) * (2 * d[0] - 7 * d[1] + 11 * d[2]);
  q.p[1] = (1.0 / 6.0) * (-d[1] + 5 * d[2] + 2 * d[3]);
  q.p[2] = (1.0 / 6.0) * (2 * d[2] + 5 * d[3] - d[4]);
  return q;
}

// Kink test: max_k(eps dx^2 + beta_k) > ratio * min_k(eps dx^2 + beta_k).
bool weno5Kink(const double *s, double dx, double ratio) {
  const double e = 1e-6 * dx * dx;
  const double a0 = e + s[0], a1 = e + s[1], a2 = e + s[2];
  const double mx = std::max(a0, std::max(a1, a2));
  const double mn = std::min(a0, std::min(a1, a2));
  return mx > ratio * mn;
}

// Network inputs of one stencil: d / max|d| (all zero when d == 0).
void mlpInputs(const double *d, float in[MLP_IN], double &scale) {
  double m = 0;
  for (int i = 0; i < 5; ++i) m = std::max(m, std::fabs(d[i]));
  scale = m;
  for (int i = 0; i < 5; ++i) in[i] = (m > 0) ? (float)(d[i] / m) : 0.f;
}

double weno5L(const double *x, double dx, bool plus, const float *P) {
  const Weno5Parts q = weno5Parts(x, plus);
  double scale;
  float in[MLP_IN];
  mlpInputs(q.d, in, scale);
  if (scale == 0. || weno5Kink(q.s, dx, MLP_KINK_RATIO)) return weno5(x, dx, plus);
  float w[MLP_OUT];
  mlpForward(P, in, w);
  return ((double)w[0] * q.p[0] + (double)w[1] * q.p[1] + (double)w[2] * q.p[2]) / dx;
}

// Mean squared loss over n samples and its gradient.  Sample s: X[8s..8s+4]
// network inputs, X[8s+5..8s+7] candidate polynomials divided by max|d|,
// T[s] target derivative times dx / max|d|.  loss = (1/n) sum (sum_k w_k
// p_k - t)^2.  emulate != 0 reproduces the rounding points of the GPU
// training kernel: weights used as their bf16 values, the inputs, both
// hidden activations and the layer-2 pre-activation gradient rounded to
// bf16 where they enter the products; everything else in R.
template <class R>
static double lossGrad(const float *Pm, int64_t n, const float *X, const float *T, double *G, int emulate) {
  std::vector<R> P(MLP_NPARAM);
  for (int i = 0; i < MLP_NPARAM; ++i) {
    const bool isW = (i < MLP_B1) || (i >= MLP_W2 && i < MLP_B2) || (i >= MLP_W3 && i < MLP_B3);
    P[i] = (emulate && isW) ? (R)bf16Round(Pm[i]) : (R)Pm[i];
  }
  auto rb = [&](R v) -> R { return emulate ? (R)bf16Round((float)v) : v; };
  std::vector<double> g(MLP_NPARAM, 0.0);
  double loss = 0;
  const R invn = (R)1 / (R)n;
  for (int64_t s = 0; s < n; ++s) {
    const float *xs = X + 8 * s;
    R x[8];
    for (int i = 0; i < 8; ++i) x[i] = (i < MLP_IN) ? rb((R)xs[i]) : (R)xs[i];
    R h1[MLP_H], d1[MLP_H], h1b[MLP_H], h2[MLP_H], d2[MLP_H], z3[MLP_OUT], w[MLP_OUT];
    for (int j = 0; j < MLP_H; ++j) {
      R z = P[MLP_B1 + j];
      for (int i = 0; i < MLP_IN; ++i) z += P[MLP_W1 + j * MLP_IN + i] * x[i];
      h1[j] = geluT<R>(z, d1[j]);
      h1b[j] = rb(h1[j]);
    }
    for (int j = 0; j < MLP_H; ++j) {
      R z = 0;
      for (int i = 0; i < MLP_H; ++i) z += P[MLP_W2 + j * MLP_H + i] * h1b[i];
      z = z + P[MLP_B2 + j];
      h2[j] = rb(geluT<R>(z, d2[j]));
    }
    for (int k = 0; k < MLP_OUT; ++k) {
      R z = P[MLP_B3 + k];
      for (int i = 0; i < MLP_H; ++i) z += P[MLP_W3 + k * MLP_H + i] * h2[i];
      z3[k] = z;
    }
    R m = std::max(z3[0], std::max(z3[1], z3[2]));
    R e[MLP_OUT], se = 0;
    for (int k = 0; k < MLP_OUT; ++k) {
      e[k] = std::exp(z3[k] - m);
      se += e[k];
    }
    for (int k = 0; k < MLP_OUT; ++k) w[k] = e[k] / se;
    const R pred = w[0] * (R)x[5] + w[1] * (R)x[6] + w[2] * (R)x[7];
    const R r = pred - (R)T[s];
    loss += (double)(r * r);
    const R dpred = (R)2 * r * invn;
    R dw[MLP_OUT], dz3[MLP_OUT];
    for (int k = 0; k < MLP_OUT; ++k) dw[k] = dpred * (R)x[5 + k];
    const R wdw = w[0] * dw[0] + w[1] * dw[1] + w[2] * dw[2];
    for (int k = 0; k < MLP_OUT; ++k) dz3[k] = w[k] * (dw[k] - wdw);
    R dz2[MLP_H], dz2b[MLP_H];
    for (int i = 0; i < MLP_H; ++i) {
      R dh = 0;
      for (int k = 0; k < MLP_OUT; ++k) dh += P[MLP_W3 + k * MLP_H + i] * dz3[k];
      dz2[i] = dh * d2[i];
      dz2b[i] = rb(dz2[i]);
    }
    for (int k = 0; k < MLP_OUT; ++k) {
      g[MLP_B3 + k] += (double)dz3[k];
      for (int i = 0; i < MLP_H; ++i) g[MLP_W3 + k * MLP_H + i] += (double)(dz3[k] * h2[i]);
    }
    for (int j = 0; j < MLP_H; ++j) {
      g[MLP_B2 + j] += (double)dz2[j];
      for (int i = 0; i < MLP_H; ++i) g[MLP_W2 + j * MLP_H + i] += (double)(dz2b[j] * h1b[i]);
    }
    for (int i = 0; i < MLP_H; ++i) {
      R dh = 0;
      for (int j = 0; j < MLP_H; ++j) dh += P[MLP_W2 + j * MLP_H + i] * dz2b[j];
      const R dz1 = dh * d1[i];
      g[MLP_B1 + i] += (double)dz1;
      for (int q = 0; q < MLP_IN; ++q) g[MLP_W1 + i * MLP_IN + q] += (double)(dz1 * (R)x[q]);
    }
  }
  for (int i = 0; i < MLP_NPARAM; ++i) G[i] = g[i];
  return loss / (double)n;
}

double mlpLossGrad(const float *P, int64_t n, const float *X, const float *T, double *G, int emulate, int useDouble) {
  if (useDouble) return lossGrad<double>(P, n, X, T, G, emulate);
  return lossGrad<float>(P, n, X, T, G, emulate);
}

}  // namespace lso

extern "C" {

void lso_mlp_forward(const float *P, const float *in, float *w) { lso::mlpForward(P, in, w); }

double lso_weno5_learned(const double *x, double dx, int plus, const float *P) {
  return lso::weno5L(x, dx, plus != 0, P);
}

double lso_mlp_loss_grad(const float *P, int64_t n, const float *X, const float *T, double *G, int emulate,
                         int useDouble) {
  return lso::mlpLossGrad(P, n, X, T, G, emulate, useDouble);
}

float lso_bf16_round(float x) { return lso::bf16Round(x); }

}  // extern "C"
