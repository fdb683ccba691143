/*
 * mlp_oracle.cpp -- CPU reference of the learned tmax proposal network.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the gfx950 MFMA kernels in
 * j3d_amd/csrc/mlp.hip.  The network has no counterpart in the reference
 * (j3d); it is specified by the workload: input = ray origin and direction
 * with 6-frequency Fourier features (78 values), three hidden layers of 64
 * (ReLU), two outputs (normalised hit distance, hit logit).
 *
 * The ray enters as the line it lies on (the hit along a line does not depend
 * on where on the line the camera sits, as long as it is outside the box):
 *   u = d (1/|d|), a = (c - o) . u (distance from o to the point of the line
 *   closest to the box centre c), o' = (o + a u - c) (1/R), x = (o', u)
 *   (fp32, evaluated in exactly this order: the GPU computes the same bits)
 *   features = [x_0..x_5, then for i = 0..5, k = 0..5: sin(2^k pi x_i), cos(2^k pi x_i)]
 *   h1 = relu(W1 f + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3), y = W4 h3 + b4
 *   loss = mean over rays of  hit * 10 * (y0 - s)^2 + softplus(y1) - hit * y1
 *   where s = (t_hit |d| - a) / R: the hit's signed distance from the line's
 *   closest point to c, in units of R (c, R: the mesh box's centre and
 *   half-diagonal).  A prediction maps back to t = (y0 R + a) / |d|.
 *
 * Parameters are one float array: W1[64][78], b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], W4[2][64], b4[2] (13,506 values).  With bf16 != 0
 * every operand that enters a matrix product on the GPU's matrix cores is
 * rounded to bfloat16 (round to nearest even): features, weights, hidden
 * activations, and in the backward pass the hidden-layer deltas (which also
 * makes the bias gradients sums of rounded deltas).  The output gradients dy
 * are not rounded: the output layer's backward runs on the vector ALUs.  Each
 * neuron's sum starts at its bias.  Sums are accumulated in double here, in
 * fp32 on the GPU, so the two agree to within fp32 summation error.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

const int kIn = 78, kH = 64, kOut = 2;
const int kParams = kH * kIn + kH + kH * kH + kH + kH * kH + kH + kOut * kH + kOut;
const float kRegWeight = 10.f;

inline float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

struct Layout {
  const float *W1, *b1, *W2, *b2, *W3, *b3, *W4, *b4;
  explicit Layout(const float* p) {
    W1 = p;
    b1 = W1 + kH * kIn;
    W2 = b1 + kH;
    b2 = W2 + kH * kH;
    W3 = b2 + kH;
    b3 = W3 + kH * kH;
    W4 = b3 + kH;
    b4 = W4 + kOut * kH;
  }
};

// The network input of a ray (fp32, evaluated in this order); returns |d|, sets a.
float line_input(const float* ray, const float* norm, float x[6], float& a) {
  const float dl = std::sqrt(ray[3] * ray[3] + ray[4] * ray[4] + ray[5] * ray[5]);
  const float rl = 1.f / dl, rR = 1.f / norm[3];
  const float u[3] = {ray[3] * rl, ray[4] * rl, ray[5] * rl};
  a = (norm[0] - ray[0]) * u[0] + (norm[1] - ray[1]) * u[1] + (norm[2] - ray[2]) * u[2];
  for (int k = 0; k < 3; ++k) x[k] = ((ray[k] + a * u[k]) - norm[k]) * rR;
  for (int k = 0; k < 3; ++k) x[3 + k] = u[k];
  return dl;
}

void features(const float* ray, const float* norm, float* f) {
  const double kPi = 3.14159265358979323846;
  float x[6], a;
  line_input(ray, norm, x, a);
  for (int i = 0; i < 6; ++i) f[i] = x[i];
  for (int i = 0; i < 6; ++i)
    for (int k = 0; k < 6; ++k) {
      double a = (double)(1 << k) * kPi * (double)x[i];
      f[6 + 12 * i + 2 * k] = (float)std::sin(a);
      f[6 + 12 * i + 2 * k + 1] = (float)std::cos(a);
    }
}

void dense(const float* W, const float* b, const float* in, int n_in, int n_out, bool relu, bool bf, float* out) {
  for (int o = 0; o < n_out; ++o) {
    double acc = b[o];
    for (int i = 0; i < n_in; ++i) {
      float w = W[o * n_in + i], a = in[i];
      if (bf) w = bf16_round(w), a = bf16_round(a);
      acc += (double)w * (double)a;
    }
    float v = (float)acc;
    out[o] = relu ? (v > 0.f ? v : 0.f) : v;
  }
}

}  // namespace

extern "C" {

int orc_mlp_param_count(void) { return kParams; }

/* Per ray: |d|, a, and the 6 network inputs x (out: n x 8). */
void orc_mlp_line(const float* rays, uint64_t n, const float* norm, float* out) {
  for (uint64_t r = 0; r < n; ++r) {
    float x[6], a;
    const float dl = line_input(rays + 8 * r, norm, x, a);
    out[8 * r] = dl;
    out[8 * r + 1] = a;
    for (int k = 0; k < 6; ++k) out[8 * r + 2 + k] = x[k];
  }
}

void orc_mlp_features(const float* rays, uint64_t n, const float* norm, float* out) {
  for (uint64_t r = 0; r < n; ++r) features(rays + 8 * r, norm, out + kIn * r);
}

/* rays: n x 8 (ox,oy,oz,dx,dy,dz,*,*); norm: center xyz, extent; out: n x 2 */
void orc_mlp_forward(const float* params, const float* norm, const float* rays, uint64_t n, int bf16, float* out) {
  Layout L(params);
  float f[kIn], h1[kH], h2[kH], h3[kH];
  for (uint64_t r = 0; r < n; ++r) {
    features(rays + 8 * r, norm, f);
    dense(L.W1, L.b1, f, kIn, kH, true, bf16 != 0, h1);
    dense(L.W2, L.b2, h1, kH, kH, true, bf16 != 0, h2);
    dense(L.W3, L.b3, h2, kH, kH, true, bf16 != 0, h3);
    dense(L.W4, L.b4, h3, kH, kOut, false, bf16 != 0, out + 2 * r);
  }
}

/* labels: n x 2 (s = normalised hit distance, hit in {0, 1}).  grads: kParams.
   Returns the mean loss. */
float orc_mlp_gradients(const float* params, const float* norm, const float* rays, const float* labels, uint64_t n,
                        int bf16, float* grads) {
  Layout L(params);
  std::vector<double> g(kParams, 0.0);
  double* gW1 = g.data();
  double* gb1 = gW1 + kH * kIn;
  double* gW2 = gb1 + kH;
  double* gb2 = gW2 + kH * kH;
  double* gW3 = gb2 + kH;
  double* gb3 = gW3 + kH * kH;
  double* gW4 = gb3 + kH;
  double* gb4 = gW4 + kOut * kH;
  const bool bf = bf16 != 0;
  auto q = [&](float x) { return bf ? bf16_round(x) : x; };
  // a hidden-layer delta as the GPU holds it: fp32, then bf16 for the matrix cores
  auto qd = [&](double x) { return bf ? (double)bf16_round((float)x) : x; };
  double loss = 0.0;
  float f[kIn], h1[kH], h2[kH], h3[kH], y[2];
  double d3[kH], d2[kH], d1[kH];
  for (uint64_t r = 0; r < n; ++r) {
    features(rays + 8 * r, norm, f);
    dense(L.W1, L.b1, f, kIn, kH, true, bf, h1);
    dense(L.W2, L.b2, h1, kH, kH, true, bf, h2);
    dense(L.W3, L.b3, h2, kH, kH, true, bf, h3);
    dense(L.W4, L.b4, h3, kH, kOut, false, bf, y);
    const float s = labels[2 * r], hit = labels[2 * r + 1];
    const double e = (double)y[0] - s;
    const double sp = std::max((double)y[1], 0.0) + std::log1p(std::exp(-std::fabs((double)y[1])));
    loss += hit * kRegWeight * e * e + sp - hit * y[1];
    const double dy0 = 2.0 * kRegWeight * hit * e / (double)n;
    const double dy1 = (1.0 / (1.0 + std::exp(-(double)y[1])) - hit) / (double)n;
    const double dy[2] = {dy0, dy1};
    for (int o = 0; o < kOut; ++o) {
      gb4[o] += dy[o];
      for (int i = 0; i < kH; ++i) gW4[o * kH + i] += dy[o] * q(h3[i]);
    }
    for (int i = 0; i < kH; ++i) {
      double a = 0.0;
      for (int o = 0; o < kOut; ++o) a += dy[o] * q(L.W4[o * kH + i]);
      d3[i] = h3[i] > 0.f ? qd(a) : 0.0;
    }
    for (int o = 0; o < kH; ++o) {
      gb3[o] += d3[o];
      for (int i = 0; i < kH; ++i) gW3[o * kH + i] += d3[o] * q(h2[i]);
    }
    for (int i = 0; i < kH; ++i) {
      double a = 0.0;
      for (int o = 0; o < kH; ++o) a += d3[o] * q(L.W3[o * kH + i]);
      d2[i] = h2[i] > 0.f ? qd(a) : 0.0;
    }
    for (int o = 0; o < kH; ++o) {
      gb2[o] += d2[o];
      for (int i = 0; i < kH; ++i) gW2[o * kH + i] += d2[o] * q(h1[i]);
    }
    for (int i = 0; i < kH; ++i) {
      double a = 0.0;
      for (int o = 0; o < kH; ++o) a += d2[o] * q(L.W2[o * kH + i]);
      d1[i] = h1[i] > 0.f ? qd(a) : 0.0;
    }
    for (int o = 0; o < kH; ++o) {
      gb1[o] += d1[o];
      for (int i = 0; i < kIn; ++i) gW1[o * kIn + i] += d1[o] * q(f[i]);
    }
  }
  for (int k = 0; k < kParams; ++k) grads[k] = (float)g[k];
  return (float)(loss / (double)n);
}

}  // extern "C"
