/*
 * feat_oracle_mlp.c -- CPU oracle of the learned smoother's network and of its
 * input features (no counterpart in the reference).  TEST INFRASTRUCTURE ONLY.
 *
 * Network (feat3amd.h): z1 = W1 x + b1, h1 = bf16(gelu(z1)), z2 = W2 h1 + b2,
 * y = w3 . gelu(z2) + b3, BF16 weights W1, W2 and inputs, FP32 arithmetic;
 * loss = mean (y - t)^2.  Sums run sequentially in index order (the device sums
 * the 64-term products inside the matrix-core instruction, so results agree to
 * FP32 rounding, not bitwise).  round_bf16 = 0 keeps every activation and
 * back-propagated delta in FP32 (the plain network, checked against float64 finite
 * differences by the tests); round_bf16 = 1 rounds exactly where the device does.
 */
#include "feat_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { IN = 54, INP = 64, HID = 64 };
enum { OW1 = 0, OB1 = 4096, OW2 = 4160, OB2 = 8256, OW3 = 8320, OB3 = 8384 };

static uint16_t f2bf(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float bf2f(uint16_t b)
{
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static float rnd(float f, int on) { return on ? bf2f(f2bf(f)) : f; }
static float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
static float gelu_d(float z)
{
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

typedef struct {
  float z1[HID], h1[HID], z2[HID], y;
} act_t;

static void forward1(const float* p, const float* w1, const float* w2, const float* x, int rb, act_t* a)
{
  for(int o = 0; o < HID; ++o)
  {
    float s = 0.0f;
    for(int i = 0; i < INP; ++i) s = fmaf(w1[o * INP + i], x[i], s);
    a->z1[o] = s + p[OB1 + o];
    a->h1[o] = rnd(gelu(a->z1[o]), rb);
  }
  for(int o = 0; o < HID; ++o)
  {
    float s = 0.0f;
    for(int i = 0; i < HID; ++i) s = fmaf(w2[o * HID + i], a->h1[i], s);
    a->z2[o] = s + p[OB2 + o];
  }
  float y = 0.0f;
  for(int o = 0; o < HID; ++o) y = fmaf(p[OW3 + o], gelu(a->z2[o]), y);
  a->y = y + p[OB3];
}

static void weights(const float* p, int rb, float* w1, float* w2)
{
  for(int i = 0; i < HID * INP; ++i) w1[i] = rnd(p[OW1 + i], rb);
  for(int i = 0; i < HID * HID; ++i) w2[i] = rnd(p[OW2 + i], rb);
}

static void input_row(const uint16_t* X, uint64_t s, float* x)
{
  for(int i = 0; i < INP; ++i) x[i] = bf2f(X[s * INP + (uint64_t)i]);
}

void orc_mlp_forward(const float* p, const uint16_t* X, uint64_t n, int round_bf16, float* y)
{
  float w1[HID * INP], w2[HID * HID], x[INP];
  act_t a;
  weights(p, round_bf16, w1, w2);
  for(uint64_t s = 0; s < n; ++s)
  {
    input_row(X, s, x);
    forward1(p, w1, w2, x, round_bf16, &a);
    y[s] = a.y;
  }
}

/* grad: FEAT_MLP_NPARAM_PAD floats, grad[last] = sum of squared errors */
void orc_mlp_grad(const float* p, const uint16_t* X, const float* t, uint64_t n, int round_bf16, float* grad)
{
  float w1[HID * INP], w2[HID * HID], x[INP], d2[HID], d1[HID];
  act_t a;
  weights(p, round_bf16, w1, w2);
  memset(grad, 0, sizeof(float) * 8448);
  const float inv_b = 1.0f / (float)n;
  float loss = 0.0f;
  for(uint64_t s = 0; s < n; ++s)
  {
    input_row(X, s, x);
    forward1(p, w1, w2, x, round_bf16, &a);
    const float d = a.y - t[s];
    loss = fmaf(d, d, loss);
    const float dy = 2.0f * d * inv_b;
    grad[OB3] += dy;
    for(int o = 0; o < HID; ++o)
    {
      const float dz2 = dy * p[OW3 + o] * gelu_d(a.z2[o]);
      grad[OW3 + o] = fmaf(dy, gelu(a.z2[o]), grad[OW3 + o]);
      grad[OB2 + o] += dz2;
      d2[o] = rnd(dz2, round_bf16);
    }
    for(int o = 0; o < HID; ++o)
      for(int i = 0; i < HID; ++i) grad[OW2 + o * HID + i] = fmaf(d2[o], a.h1[i], grad[OW2 + o * HID + i]);
    for(int i = 0; i < HID; ++i)
    {
      float dh = 0.0f;
      for(int o = 0; o < HID; ++o) dh = fmaf(w2[o * HID + i], d2[o], dh);
      const float dz1 = dh * gelu_d(a.z1[i]);
      grad[OB1 + i] += dz1;
      d1[i] = rnd(dz1, round_bf16);
    }
    for(int o = 0; o < HID; ++o)
      for(int i = 0; i < IN; ++i) grad[OW1 + o * INP + i] = fmaf(d1[o], x[i], grad[OW1 + o * INP + i]);
  }
  grad[8447] = loss;
}

/* inputs of interior node i on an m^3 grid (the device's node_features) */
static int features(uint64_t m, uint64_t i, const double* r, const double* dinv, const float* cfeat, uint16_t* xrow,
                    double* s_out)
{
  const uint64_t x = i % m, y = (i / m) % m, z = i / (m * m);
  double rho[27], s = 0.0;
  const int interior = x > 0 && y > 0 && z > 0 && x + 1 < m && y + 1 < m && z + 1 < m;
  if(interior)
  {
    int j = 0;
    for(int dz = -1; dz <= 1; ++dz)
      for(int dy = -1; dy <= 1; ++dy)
        for(int dx = -1; dx <= 1; ++dx, ++j)
        {
          const uint64_t nb = (uint64_t)((int64_t)i + dx + dy * (int64_t)m + dz * (int64_t)(m * m));
          rho[j] = r[nb] * dinv[i];
          s = fmax(s, fabs(rho[j]));
        }
  }
  *s_out = s;
  if(!interior || s == 0.0)
  {
    memset(xrow, 0, sizeof(uint16_t) * INP);
    return 0;
  }
  int j = 0;
  for(int dz = -1; dz <= 1; ++dz)
    for(int dy = -1; dy <= 1; ++dy)
      for(int dx = -1; dx <= 1; ++dx, ++j)
      {
        const uint64_t nb = (uint64_t)((int64_t)i + dx + dy * (int64_t)m + dz * (int64_t)(m * m));
        xrow[j] = f2bf((float)(rho[j] / s));
        xrow[27 + j] = f2bf(cfeat[nb]);
      }
  for(int k = IN; k < INP; ++k) xrow[k] = 0;
  return 1;
}

void orc_mlp_features(uint64_t m, const double* r, const double* e, const double* dinv, const float* cfeat,
                      const uint64_t* node, uint64_t nsamp, uint16_t* X, float* t)
{
  for(uint64_t k = 0; k < nsamp; ++k)
  {
    double s;
    const int ok = features(m, node[k], r, dinv, cfeat, X + k * INP, &s);
    t[k] = ok ? (float)(e[node[k]] / s) : 0.0f;
  }
}

/* x_i += omega * (jacobi * r_i / a_ii + s_i * y_i) on every interior node (the device's learned smoothing step) */
void orc_mlp_smooth(uint64_t m, const float* p, const double* r, const double* dinv, const float* cfeat, double omega,
                    double jacobi, double* x)
{
  const uint64_t n = m * m * m;
  float w1[HID * INP], w2[HID * HID], xin[INP];
  uint16_t row[INP];
  act_t a;
  weights(p, 1, w1, w2);
  double* upd = (double*)calloc(n, sizeof(double));
  char* has = (char*)calloc(n, 1);
  for(uint64_t i = 0; i < n; ++i)
  {
    double s;
    if(!features(m, i, r, dinv, cfeat, row, &s)) continue;
    for(int k = 0; k < INP; ++k) xin[k] = bf2f(row[k]);
    forward1(p, w1, w2, xin, 1, &a);
    upd[i] = fma(jacobi, r[i] * dinv[i], s * (double)a.y);
    has[i] = 1;
  }
  for(uint64_t i = 0; i < n; ++i)
    if(has[i]) x[i] = fma(omega, upd[i], x[i]);
  free(upd);
  free(has);
}
