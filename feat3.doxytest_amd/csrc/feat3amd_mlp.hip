// feat3amd_mlp.hip -- the learned smoother's network on gfx950 matrix cores.
//
// MLP 54 -> 64 -> 64 -> 1 with GELU: inputs padded to 64, BF16 weights and
// activations, FP32 accumulation.  Layers 1 and 2 (and, in training, the three
// 64x64 products of the backward pass) run on v_mfma_f32_16x16x16_bf16: one
// wavefront owns 16 samples; a workgroup of four wavefronts owns 64.  Operands
// are staged in LDS; the 16x16 MFMA lane layout is
//   A: lane l holds A[l%16][k0 + 4*(l/16) + i], B: B[k0 + 4*(l/16) + i][l%16],
//   D: lane l holds D[4*(l/16) + i][l%16], i = 0..3.
// Per-sample and per-unit sums outside the MFMAs (layer 3, bias and layer-3
// gradients, the loss) run sequentially in a fixed order, and weight gradients
// are reduced over wavefronts in index order: training is deterministic.
#include "feat3amd_common.hpp"

namespace {

constexpr int kIn = FEAT_MLP_IN, kInP = 64, kHid = FEAT_MLP_HID;
constexpr int kW1 = 0, kB1 = 4096, kW2 = 4160, kB2 = 8256, kW3 = 8320, kB3 = 8384;
constexpr int kLoss = FEAT_MLP_NPARAM_PAD - 1;
constexpr int kWaves = 4;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
  // round to nearest even: gfx950's v_cvt_pk_bf16_f32
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
// GELU of the inference kernels: erf by Abramowitz & Stegun 7.1.26 (|error| <= 1.5e-7,
// far below the BF16 rounding its result goes through), one reciprocal and one exp
__device__ __forceinline__ float gelu_fast(float z)
{
  const float x = fabsf(z) * 0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(__fmaf_rn(0.3275911f, x, 1.0f));
  float p = 1.061405429f;
  p = __fmaf_rn(p, t, -1.453152027f);
  p = __fmaf_rn(p, t, 1.421413741f);
  p = __fmaf_rn(p, t, -0.284496736f);
  p = __fmaf_rn(p, t, 0.254829592f);
  const float e = 1.0f - p * t * __expf(-x * x);
  return 0.5f * z * (1.0f + copysignf(e, z));
}
__device__ __forceinline__ float gelu_d(float z)
{
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// acc += A(16 x K) * B(K x 16), A(m,k) = A[m*ars + k*acs], B(k,n) = B[k*brs + n*bcs], BF16 in LDS
__device__ __forceinline__ f32x4 mfma_tile(f32x4 acc, const uint16_t* A, int ars, int acs, const uint16_t* B, int brs,
                                           int bcs, int K)
{
  const int l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  for(int k0 = 0; k0 < K; k0 += 16)
  {
    bf16x4 a, b;
#pragma unroll
    for(int i = 0; i < 4; ++i)
    {
      a[i] = short(A[r * ars + (k0 + kq + i) * acs]);
      b[i] = short(B[(k0 + kq + i) * brs + r * bcs]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

struct SharedWeights
{
  uint16_t w1[kHid * kInP];
  uint16_t w2[kHid * kHid];
  float b1[kHid], b2[kHid], w3[kHid];
  float b3;
};

__device__ void load_weights(SharedWeights& W, const float* __restrict__ p, const uint16_t* __restrict__ wbf)
{
  for(int i = threadIdx.x; i < kHid * kInP; i += blockDim.x)
  {
    W.w1[i] = wbf[i];
    W.w2[i] = wbf[kHid * kInP + i];
  }
  for(int i = threadIdx.x; i < kHid; i += blockDim.x)
  {
    W.b1[i] = p[kB1 + i];
    W.b2[i] = p[kB2 + i];
    W.w3[i] = p[kW3 + i];
  }
  if(threadIdx.x == 0) W.b3 = p[kB3];
}

struct WaveBuf
{
  uint16_t x[16 * kInP];
  uint16_t h1[16 * kHid];
  float z1[16 * kHid];
  float z2[16 * kHid];
  uint16_t d2[16 * kHid];
  uint16_t d1[16 * kHid];
  float y[16];
};

// layers 1 and 2 for the wave's 16 samples in B.x: fills z1, h1, z2 (LDS); returns nothing
__device__ void forward12(const SharedWeights& W, WaveBuf& B)
{
  const int l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  for(int nt = 0; nt < 4; ++nt)
  {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(acc, B.x, kInP, 1, W.w1 + nt * 16 * kInP, 1, kInP, kInP);
#pragma unroll
    for(int i = 0; i < 4; ++i) B.z1[(kq + i) * kHid + nt * 16 + r] = acc[i] + W.b1[nt * 16 + r];
  }
  __syncthreads();
  for(int e = l; e < 16 * kHid; e += 64) B.h1[e] = f2bf(gelu(B.z1[e]));
  __syncthreads();
  for(int nt = 0; nt < 4; ++nt)
  {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(acc, B.h1, kHid, 1, W.w2 + nt * 16 * kHid, 1, kHid, kHid);
#pragma unroll
    for(int i = 0; i < 4; ++i) B.z2[(kq + i) * kHid + nt * 16 + r] = acc[i] + W.b2[nt * 16 + r];
  }
  __syncthreads();
  // layer 3 in FP32: y = b3 + sum_o w3[o] gelu(z2[o]), sequential in o
  if(l < 16)
  {
    float y = 0.0f;
    for(int o = 0; o < kHid; ++o) y = __fmaf_rn(W.w3[o], gelu(B.z2[l * kHid + o]), y);
    B.y[l] = y + W.b3;
  }
  __syncthreads();
}

// Training: forward + backward over tiles of 64 samples per workgroup; per-wave
// gradient partials (FEAT_MLP_NPARAM_PAD floats each, loss in the last slot).
__global__ __launch_bounds__(64 * kWaves) void k_mlp_train(const float* __restrict__ p, const uint16_t* __restrict__ wbf,
                                                           const uint16_t* __restrict__ X, const float* __restrict__ t,
                                                           uint64_t nsamp, float inv_b, float* __restrict__ partials)
{
  __shared__ SharedWeights W;
  __shared__ WaveBuf Bw[kWaves];
  __shared__ float gb1[kWaves][kHid], gb2[kWaves][kHid], gw3[kWaves][kHid], gb3[kWaves], gloss[kWaves];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
  WaveBuf& B = Bw[wave];
  load_weights(W, p, wbf);
  gb1[wave][l] = gb2[wave][l] = gw3[wave][l] = 0.0f;
  if(l == 0) gb3[wave] = gloss[wave] = 0.0f;
  f32x4 aw1[4][4], aw2[4][4];
  for(int a = 0; a < 4; ++a)
    for(int b = 0; b < 4; ++b) aw1[a][b] = aw2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  __syncthreads();
  const uint64_t ngroups = (nsamp + 16 * kWaves - 1) / (16 * kWaves);
  for(uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x)
  {
    const uint64_t s0 = (g * kWaves + uint64_t(wave)) * 16;
    for(int e = l; e < 16 * kInP; e += 64)
    {
      const uint64_t s = s0 + uint64_t(e / kInP);
      B.x[e] = s < nsamp ? X[s * kInP + uint64_t(e % kInP)] : uint16_t(0);
    }
    __syncthreads();
    forward12(W, B);
    // dy, loss, layer-3 gradients: lane o walks the 16 samples in order
    __shared__ float dyv[kWaves][16];
    if(l < 16)
    {
      const uint64_t s = s0 + uint64_t(l);
      float dy = 0.0f;
      if(s < nsamp)
      {
        const float d = B.y[l] - t[s];
        dy = 2.0f * d * inv_b;
      }
      dyv[wave][l] = dy;
    }
    __syncthreads();
    if(l == 0)
      for(int s = 0; s < 16; ++s)
        if(s0 + uint64_t(s) < nsamp)
        {
          const float d = B.y[s] - t[s0 + uint64_t(s)];
          gloss[wave] = __fmaf_rn(d, d, gloss[wave]);
          gb3[wave] += dyv[wave][s];
        }
    {
      const int o = l;
      float a2 = gb2[wave][o], a3 = gw3[wave][o];
      for(int s = 0; s < 16; ++s)
      {
        const float z = B.z2[s * kHid + o], dy = dyv[wave][s];
        const float dz2 = dy * W.w3[o] * gelu_d(z);
        a3 = __fmaf_rn(dy, gelu(z), a3);
        a2 += dz2;
        B.d2[s * kHid + o] = f2bf(dz2);
      }
      gb2[wave][o] = a2;
      gw3[wave][o] = a3;
    }
    __syncthreads();
    // dW2 += dZ2^T H1 (M = out unit, N = in unit, K = samples)
    for(int mt = 0; mt < 4; ++mt)
      for(int nt = 0; nt < 4; ++nt)
        aw2[mt][nt] = mfma_tile(aw2[mt][nt], B.d2 + mt * 16, 1, kHid, B.h1 + nt * 16, kHid, 1, 16);
    // dH1 = dZ2 W2, dZ1 = dH1 * gelu'(Z1) (stored FP32 in z2's place, BF16 in d1)
    for(int nt = 0; nt < 4; ++nt)
    {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma_tile(acc, B.d2, kHid, 1, W.w2 + nt * 16, kHid, 1, kHid);
#pragma unroll
      for(int i = 0; i < 4; ++i)
      {
        const int e = (kq + i) * kHid + nt * 16 + r;
        const float dz1 = acc[i] * gelu_d(B.z1[e]);
        B.z2[e] = dz1;
        B.d1[e] = f2bf(dz1);
      }
    }
    __syncthreads();
    {
      const int o = l;
      float a1 = gb1[wave][o];
      for(int s = 0; s < 16; ++s) a1 += B.z2[s * kHid + o];
      gb1[wave][o] = a1;
    }
    // dW1 += dZ1^T X
    for(int mt = 0; mt < 4; ++mt)
      for(int nt = 0; nt < 4; ++nt)
        aw1[mt][nt] = mfma_tile(aw1[mt][nt], B.d1 + mt * 16, 1, kHid, B.x + nt * 16, kInP, 1, 16);
    __syncthreads();
  }
  float* out = partials + (uint64_t(blockIdx.x) * kWaves + uint64_t(wave)) * FEAT_MLP_NPARAM_PAD;
  for(int mt = 0; mt < 4; ++mt)
    for(int nt = 0; nt < 4; ++nt)
#pragma unroll
      for(int i = 0; i < 4; ++i)
      {
        const int row = mt * 16 + kq + i, col = nt * 16 + r;
        out[kW1 + row * kInP + col] = aw1[mt][nt][i];
        out[kW2 + row * kHid + col] = aw2[mt][nt][i];
      }
  out[kB1 + l] = gb1[wave][l];
  out[kB2 + l] = gb2[wave][l];
  out[kW3 + l] = gw3[wave][l];
  if(l == 0)
  {
    out[kB3] = gb3[wave];
    out[kLoss] = gloss[wave];
  }
  for(int i = kB3 + 1 + l; i < kLoss; i += 64) out[i] = 0.0f;
}

// Gradient reduction in a fixed order: segment g sums partial vectors 32g .. 32g+31 in
// index order (stage 0), then grad[i] sums the segments in index order (stage 1);
// padded inputs get no gradient.
constexpr int kRedSeg = 32;
__global__ void k_mlp_reduce_seg(const float* __restrict__ partials, uint64_t nparts, float* __restrict__ seg)
{
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= uint64_t(FEAT_MLP_NPARAM_PAD)) return;
  const uint64_t w0 = uint64_t(blockIdx.y) * kRedSeg, w1 = w0 + kRedSeg < nparts ? w0 + kRedSeg : nparts;
  float s = 0.0f;
  for(uint64_t w = w0; w < w1; ++w) s += partials[w * FEAT_MLP_NPARAM_PAD + i];
  seg[uint64_t(blockIdx.y) * FEAT_MLP_NPARAM_PAD + i] = s;
}
__global__ void k_mlp_reduce(const float* __restrict__ seg, uint64_t nseg, float* __restrict__ grad)
{
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= uint64_t(FEAT_MLP_NPARAM_PAD)) return;
  float s = 0.0f;
  for(uint64_t w = 0; w < nseg; ++w) s += seg[w * FEAT_MLP_NPARAM_PAD + i];
  if(i < uint64_t(kB1) && (i % kInP) >= uint64_t(kIn)) s = 0.0f;
  grad[i] = s;
}

__global__ void k_mlp_adam(float* __restrict__ p, uint16_t* __restrict__ wbf, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
                           float c1, float c2)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= FEAT_MLP_NPARAM) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float w = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  if(i < kB1 && (i % kInP) >= kIn) w = 0.0f;
  p[i] = w;
  if(i < kB1)
    wbf[i] = f2bf(w);
  else if(i >= kW2 && i < kB2)
    wbf[kHid * kInP + (i - kW2)] = f2bf(w);
}

__global__ void k_mlp_bf16(const float* __restrict__ p, uint16_t* __restrict__ wbf)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < kHid * kInP) wbf[i] = f2bf(p[kW1 + i]);
  else if(i < 2 * kHid * kInP) wbf[i] = f2bf(p[kW2 + (i - kHid * kInP)]);
}

// Weights of the inference kernels, staged once per workgroup in LDS: the MFMA A-operand
// fragments of W1 and W2 (fragment (nt, ks) of lane l = W[nt*16 + l%16][ks*16 + 4(l/16) + i],
// i = 0..3, one 8-byte read per MFMA) and the biases / output weights.
struct LdsWeights
{
  bf16x4 w1[16 * 64], w2[16 * 64];
  float b1[kHid], b2[kHid], w3[kHid];
  float b3;
};

__device__ void load_lds_weights(LdsWeights& W, const float* __restrict__ p, const uint16_t* __restrict__ wbf)
{
  for(int f = threadIdx.x; f < 16 * 64; f += blockDim.x)
  {
    const int ntks = f >> 6, l = f & 63, nt = ntks >> 2, ks = ntks & 3, c = l & 15, q = l >> 4;
    W.w1[f] = *reinterpret_cast<const bf16x4*>(wbf + (nt * 16 + c) * kInP + ks * 16 + 4 * q);
    W.w2[f] = *reinterpret_cast<const bf16x4*>(wbf + kHid * kInP + (nt * 16 + c) * kHid + ks * 16 + 4 * q);
  }
  for(int i = threadIdx.x; i < kHid; i += blockDim.x)
  {
    W.b1[i] = p[kB1 + i];
    W.b2[i] = p[kB2 + i];
    W.w3[i] = p[kW3 + i];
  }
  if(threadIdx.x == 0) W.b3 = p[kB3];
  __syncthreads();
}

// Forward pass of the wavefront's 16 samples.  X: the transposed-input operand, lane l
// holding inputs ks*16 + 4(l/16) + i of sample l%16.  Each layer is computed transposed,
// Z^T = W * X^T (weights as the A operand), so its output arrives in exactly the operand
// layout the next layer needs: lane l holds units nt*16 + 4(l/16) + i of sample l%16, and
// no transpose through memory is needed.  Returns y of sample l%16 in all four lanes of
// that sample (layer 3: the lane's 16 units in order, then over the four lanes).
__device__ __forceinline__ float forward_t(const LdsWeights& W, const bf16x4 (&X)[4])
{
  const int l = threadIdx.x & 63, q = l >> 4;
  bf16x4 H[4];
#pragma unroll
  for(int nt = 0; nt < 4; ++nt)
  {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for(int ks = 0; ks < 4; ++ks)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W.w1[(nt * 4 + ks) * 64 + l], X[ks], acc, 0, 0, 0);
#pragma unroll
    for(int i = 0; i < 4; ++i) H[nt][i] = short(f2bf(gelu_fast(acc[i] + W.b1[nt * 16 + 4 * q + i])));
  }
  float part = 0.0f;
#pragma unroll
  for(int nt = 0; nt < 4; ++nt)
  {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for(int ks = 0; ks < 4; ++ks)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W.w2[(nt * 4 + ks) * 64 + l], H[ks], acc, 0, 0, 0);
#pragma unroll
    for(int i = 0; i < 4; ++i)
    {
      const int u = nt * 16 + 4 * q + i;
      part = __fmaf_rn(W.w3[u], gelu_fast(acc[i] + W.b2[u]), part);
    }
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  return part + W.b3;
}

// Plain inference on given inputs (training, tests, evaluation): each lane loads its
// 4 x 4 inputs of sample l%16 straight into the operand (8-byte loads of the 128-byte rows).
__global__ __launch_bounds__(64 * kWaves) void k_mlp_infer(const float* __restrict__ p, const uint16_t* __restrict__ wbf,
                                                           const uint16_t* __restrict__ X, uint64_t nsamp,
                                                           float* __restrict__ y)
{
  __shared__ LdsWeights W;
  load_lds_weights(W, p, wbf);
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  const uint64_t ngroups = (nsamp + 16 * kWaves - 1) / (16 * kWaves);
  for(uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x)
  {
    asm volatile("" ::: "memory");
    const uint64_t smp = (g * kWaves + uint64_t(wave)) * 16 + uint64_t(c);
    bf16x4 Xo[4];
#pragma unroll
    for(int ks = 0; ks < 4; ++ks)
    {
      const bf16x4 v = *reinterpret_cast<const bf16x4*>(X + (smp < nsamp ? smp : 0) * kInP + ks * 16 + 4 * q);
      Xo[ks] = smp < nsamp ? v : bf16x4{0, 0, 0, 0};
    }
    const float yv = forward_t(W, Xo);
    if(q == 0 && smp < nsamp) y[smp] = yv;
  }
}

struct Grid3
{
  uint64_t m, n;
};

// features of node i from the residual r (and the scale s = max |r_j / a_ii| over the patch)
__device__ bool node_features(const Grid3& G, uint64_t i, const double* __restrict__ r, const double* __restrict__ dinv,
                              const float* __restrict__ cfeat, uint16_t* xrow, double& s)
{
  const uint64_t m = G.m;
  const uint64_t x = i % m, y = (i / m) % m, z = i / (m * m);
  const bool interior = x > 0 && y > 0 && z > 0 && x + 1 < m && y + 1 < m && z + 1 < m;
  double rho[27];
  s = 0.0;
  if(interior)
  {
    const double di = dinv[i];
    int j = 0;
    for(int dz = -1; dz <= 1; ++dz)
      for(int dy = -1; dy <= 1; ++dy)
        for(int dx = -1; dx <= 1; ++dx, ++j)
        {
          const uint64_t nb = uint64_t(int64_t(i) + dx + dy * int64_t(m) + dz * int64_t(m * m));
          rho[j] = r[nb] * di;
          s = fmax(s, fabs(rho[j]));
        }
  }
  if(!interior || s == 0.0)
  {
    for(int k = 0; k < kInP; ++k) xrow[k] = 0;
    return false;
  }
  int j = 0;
  for(int dz = -1; dz <= 1; ++dz)
    for(int dy = -1; dy <= 1; ++dy)
      for(int dx = -1; dx <= 1; ++dx, ++j)
      {
        const uint64_t nb = uint64_t(int64_t(i) + dx + dy * int64_t(m) + dz * int64_t(m * m));
        xrow[j] = f2bf(float(rho[j] / s));
        xrow[27 + j] = f2bf(cfeat[nb]);
      }
  for(int k = kIn; k < kInP; ++k) xrow[k] = 0;
  return true;
}

// training samples: node[k] of snapshot snap[k]; X row and target e_i / s
__global__ void k_mlp_features(Grid3 G, const double* __restrict__ R, const double* __restrict__ E,
                               const double* __restrict__ dinv, const float* __restrict__ cfeat,
                               const uint64_t* __restrict__ node, const uint32_t* __restrict__ snap, uint64_t nsamp,
                               uint16_t* __restrict__ X, float* __restrict__ t)
{
  const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(k >= nsamp) return;
  const uint64_t i = node[k];
  const double* r = R + uint64_t(snap[k]) * G.n;
  const double* e = E + uint64_t(snap[k]) * G.n;
  double s;
  const bool ok = node_features(G, i, r, dinv, cfeat, X + k * kInP, s);
  t[k] = ok ? float(e[i] / s) : 0.0f;
}


// The learned smoothing direction on the interior nodes,
//   c_i = omega * (jacobi * r_i / a_ii + s_i * f(patch_i)),
// written to out (DIR: c, 0 on boundary nodes and where s_i = 0, and the dot c . r is
// reduced) or added to it (x += c).  Each wavefront takes 16 nodes; lane l builds 16 of
// the 64 inputs of node l%16 (quarter l/16) in the operand layout, the four lanes of a
// node combine their maxima for s, and the lanes of quarter 0 write the result.
template<bool DIR>
__global__ __launch_bounds__(64 * kWaves) void k_mlp_smooth(Grid3 G, const float* __restrict__ p,
                                                            const uint16_t* __restrict__ wbf,
                                                            const double* __restrict__ r, const double* __restrict__ dinv,
                                                            const float* __restrict__ cfeat, double omega, double jacobi,
                                                            double* __restrict__ out, GridReduce gr)
{
  __shared__ LdsWeights W;
  load_lds_weights(W, p, wbf);
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  const uint64_t m = G.m, mm = m * m;
  const uint64_t ngroups = (G.n + 16 * kWaves - 1) / (16 * kWaves);
  double dot = 0.0;
  // patch offsets of this lane's 16 inputs (k = ks*16 + 4q + i; residual k < 27, coefficient
  // 27 <= k < 54 at patch position k - 27; padding reads the node itself)
  int32_t poff[16];
#pragma unroll
  for(int e = 0; e < 16; ++e)
  {
    const int k = (e >> 2) * 16 + 4 * q + (e & 3);
    const int j = k < 27 ? k : (k < kIn ? k - 27 : 13);
    poff[e] = int32_t(j % 3 - 1) + int32_t((j / 3) % 3 - 1) * int32_t(m) + int32_t(j / 9 - 1) * int32_t(mm);
  }
  for(uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x)
  {
    // keeps the weight fragments in LDS (read per use) instead of 64+ registers
    asm volatile("" ::: "memory");
    const uint64_t node = (g * kWaves + uint64_t(wave)) * 16 + uint64_t(c);
    const bool valid = node < G.n;
    // node indices fit 32 bits (m^3 < 2^32 is checked at launch)
    const uint32_t n32 = uint32_t(node), m32 = uint32_t(m), mm32 = uint32_t(mm);
    bool interior = false;
    if(valid)
    {
      const uint32_t xx = n32 % m32, yy = (n32 / m32) % m32, zz = n32 / mm32;
      interior = xx > 0 && yy > 0 && zz > 0 && xx + 1 < m32 && yy + 1 < m32 && zz + 1 < m32;
    }
    // every load is issued before any is used (no branches around them): patch
    // entries this lane does not need read element 0, which always exists
    const uint32_t nd = valid ? n32 : 0u;
    const double di = dinv[nd];
    const double r_c = r[nd];
    const double o_c = DIR ? 0.0 : out[nd];
    double rv[16];
    float fv[16];
#pragma unroll
    for(int ks = 0; ks < 4; ++ks)
#pragma unroll
      for(int i = 0; i < 4; ++i)
      {
        const int k = ks * 16 + 4 * q + i, e = ks * 4 + i;
        const uint32_t nb = interior ? uint32_t(int32_t(n32) + poff[e]) : 0u;
        rv[e] = r[k < 27 ? nb : 0u];
        fv[e] = cfeat[(k >= 27 && k < kIn) ? nb : 0u];
      }
    double mx = 0.0;
#pragma unroll
    for(int ks = 0; ks < 4; ++ks)
#pragma unroll
      for(int i = 0; i < 4; ++i)
      {
        const int k = ks * 16 + 4 * q + i, e = ks * 4 + i;
        rv[e] = (interior && k < 27) ? rv[e] * di : 0.0;
        fv[e] = (interior && k >= 27 && k < kIn) ? fv[e] : 0.0f;
        mx = fmax(mx, fabs(rv[e]));
      }
    // s = max over the node's 27 residuals: the four lanes c, c+16, c+32, c+48
    mx = fmax(mx, __shfl_xor(mx, 16, 64));
    mx = fmax(mx, __shfl_xor(mx, 32, 64));
    const double s = mx;
    const double inv_s = s > 0.0 ? 1.0 / s : 0.0;
    bf16x4 Xo[4];
#pragma unroll
    for(int ks = 0; ks < 4; ++ks)
#pragma unroll
      for(int i = 0; i < 4; ++i)
      {
        const int k = ks * 16 + 4 * q + i, e = ks * 4 + i;
        float v = 0.0f;
        if(s > 0.0) v = k < 27 ? float(rv[e] * inv_s) : fv[e];
        Xo[ks][i] = short(f2bf(v));
      }
    const float y = forward_t(W, Xo);
    if(q == 0 && valid)
    {
      const bool act = interior && s > 0.0;
      const double cval = act ? __fma_rn(omega, __fma_rn(jacobi, r_c * di, s * double(y)), 0.0) : 0.0;
      if(DIR)
      {
        out[node] = cval;
        dot = __fma_rn(cval, r_c, dot);
      }
      else if(act)
        out[node] = __fma_rn(omega, __fma_rn(jacobi, r_c * di, s * double(y)), o_c);
    }
  }
  if(DIR) grid_reduce_finish(gr, dot, 0.0);
}

// inputs of every node of the grid (training on whole fields): X rows, s, rho_i = r_i / a_ii
__global__ void k_mlp_features_full(Grid3 G, const double* __restrict__ r, const double* __restrict__ dinv,
                                    const float* __restrict__ cfeat, uint16_t* __restrict__ X, double* __restrict__ sc,
                                    double* __restrict__ rho)
{
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= G.n) return;
  double s;
  if(!node_features(G, i, r, dinv, cfeat, X + i * kInP, s)) s = 0.0;
  sc[i] = s;
  rho[i] = r[i] * dinv[i];
}

// stage 0: e' = e - (coef rho + s y) where s > 0, else 0 (e == nullptr: e = 0)
// stage 1: t = y + coef s q / 2 (the target whose squared-error gradient is the energy gradient)
__global__ void k_mlp_energy(uint64_t n, int stage, double coef, const float* __restrict__ y,
                             const double* __restrict__ sc, const double* __restrict__ rho, const double* __restrict__ e,
                             const double* __restrict__ q, double* __restrict__ ep, float* __restrict__ t)
{
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double s = sc[i];
  if(stage == 0)
    ep[i] = s > 0.0 ? (e ? e[i] : 0.0) - __fma_rn(coef, rho[i], s * double(y[i])) : 0.0;
  else
    t[i] = s > 0.0 ? float(double(y[i]) + 0.5 * coef * s * q[i]) : y[i];
}

// energy-optimal step along c: alpha = a / b (0 unless b > 0 and the ratio is finite),
// x += alpha c, d -= alpha q
__global__ void k_mlp_step(uint64_t n, double* __restrict__ x, double* __restrict__ d, const double* __restrict__ c,
                           const double* __restrict__ q, const double* __restrict__ a, const double* __restrict__ b)
{
  const double bb = *b, al = bb > 0.0 ? (*a) / bb : 0.0;
  const double alpha = std::isfinite(al) ? al : 0.0;
  for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
  {
    x[i] = __fma_rn(alpha, c[i], x[i]);
    d[i] = __fma_rn(-alpha, q[i], d[i]);
  }
}

int mlp_grid(uint64_t nsamp)
{
  const uint64_t g = (nsamp + 64 * kWaves - 1) / (64 * kWaves);
  return int(g < 2048 ? (g ? g : 1) : 2048);
}

} // namespace

extern "C" {

int feat_mlp_to_bf16(feat_handle_t h, const float* params, uint16_t* wbf)
{
  if(!h || !params || !wbf) return fail(FEAT_ERR_INVALID, "feat_mlp_to_bf16: null argument");
  hipLaunchKernelGGL(k_mlp_bf16, dim3((2 * kHid * kInP + 255) / 256), dim3(256), 0, h->stream, params, wbf);
  return check_launch("feat_mlp_to_bf16");
}

int feat_mlp_infer(feat_handle_t h, const float* params, const uint16_t* wbf, const uint16_t* X, uint64_t nsamp,
                   float* y)
{
  if(!h || !params || !wbf || (nsamp && (!X || !y))) return fail(FEAT_ERR_INVALID, "feat_mlp_infer: null argument");
  if(!nsamp) return FEAT_OK;
  hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_grid(nsamp)), dim3(64 * kWaves), 0, h->stream, params, wbf, X, nsamp, y);
  return check_launch("feat_mlp_infer");
}

int feat_mlp_grad(feat_handle_t h, const float* params, const uint16_t* wbf, const uint16_t* X, const float* t,
                  uint64_t nsamp, float* partials, uint64_t partials_cap, float* grad)
{
  if(!h || !params || !wbf || !X || !t || !partials || !grad || !nsamp)
    return fail(FEAT_ERR_INVALID, "feat_mlp_grad: bad argument");
  // at most 256 workgroups (each strides over its sample groups): few partial vectors
  const uint64_t ng = (nsamp + 16 * kWaves - 1) / (16 * kWaves);
  const int grid = int(ng < 256 ? ng : 256);
  const uint64_t nparts = uint64_t(grid) * kWaves, nseg = (nparts + kRedSeg - 1) / kRedSeg;
  if(nparts + nseg > partials_cap) return fail(FEAT_ERR_INVALID, "feat_mlp_grad: partials buffer too small");
  float* seg = partials + nparts * FEAT_MLP_NPARAM_PAD;
  hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(64 * kWaves), 0, h->stream, params, wbf, X, t, nsamp,
                     1.0f / float(nsamp), partials);
  hipLaunchKernelGGL(k_mlp_reduce_seg, dim3((FEAT_MLP_NPARAM_PAD + 255) / 256, unsigned(nseg)), dim3(256), 0, h->stream,
                     partials, nparts, seg);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((FEAT_MLP_NPARAM_PAD + 255) / 256), dim3(256), 0, h->stream, seg, nseg, grad);
  return check_launch("feat_mlp_grad");
}

int feat_mlp_adam(feat_handle_t h, float* params, uint16_t* wbf, const float* grad, float* m, float* v, int step,
                  float lr, float beta1, float beta2, float eps)
{
  if(!h || !params || !wbf || !grad || !m || !v || step < 1) return fail(FEAT_ERR_INVALID, "feat_mlp_adam: bad argument");
  const float c1 = 1.0f - powf(beta1, float(step)), c2 = 1.0f - powf(beta2, float(step));
  hipLaunchKernelGGL(k_mlp_adam, dim3((FEAT_MLP_NPARAM + 255) / 256), dim3(256), 0, h->stream, params, wbf, grad, m, v,
                     lr, beta1, beta2, eps, c1, c2);
  return check_launch("feat_mlp_adam");
}

int feat_mlp_features(feat_handle_t h, uint64_t m, const double* R, const double* E, const double* dinv,
                      const float* cfeat, const uint64_t* node, const uint32_t* snap, uint64_t nsamp, uint16_t* X,
                      float* t)
{
  if(!h || m < 3 || !R || !E || !dinv || !cfeat || (nsamp && (!node || !snap || !X || !t)))
    return fail(FEAT_ERR_INVALID, "feat_mlp_features: bad argument");
  if(!nsamp) return FEAT_OK;
  Grid3 G{m, m * m * m};
  hipLaunchKernelGGL(k_mlp_features, dim3((nsamp + 255) / 256), dim3(256), 0, h->stream, G, R, E, dinv, cfeat, node, snap,
                     nsamp, X, t);
  return check_launch("feat_mlp_features");
}

int feat_mlp_smooth_f64(feat_handle_t h, uint64_t m, const float* params, const uint16_t* wbf, const double* r,
                        const double* dinv, const float* cfeat, double omega, double jacobi, double* x)
{
  if(!h || m < 3 || m > 1600 || !params || !wbf || !r || !dinv || !cfeat || !x)
    return fail(FEAT_ERR_INVALID, "feat_mlp_smooth_f64: bad argument");
  if(x == r) return fail(FEAT_ERR_ALIAS, "feat_mlp_smooth_f64: x and r must differ");
  Grid3 G{m, m * m * m};
  GridReduce none{nullptr, nullptr, nullptr, nullptr, 0};
  hipLaunchKernelGGL(k_mlp_smooth<false>, dim3(mlp_grid(G.n)), dim3(64 * kWaves), 0, h->stream, G, params, wbf, r, dinv,
                     cfeat, omega, jacobi, x, none);
  return check_launch("feat_mlp_smooth_f64");
}

int feat_mlp_direction_f64(feat_handle_t h, uint64_t m, const float* params, const uint16_t* wbf, const double* r,
                           const double* dinv, const float* cfeat, double omega, double jacobi, double* c, double* dot,
                           const feat_reduce_ws_t* ws)
{
  if(!h || m < 3 || m > 1600 || !params || !wbf || !r || !dinv || !cfeat || !c || !dot)
    return fail(FEAT_ERR_INVALID, "feat_mlp_direction_f64: bad argument");
  if(c == r) return fail(FEAT_ERR_ALIAS, "feat_mlp_direction_f64: c and r must differ");
  if(need_ws(ws, dot)) return fail(FEAT_ERR_INVALID, "feat_mlp_direction_f64: reduction workspace missing");
  Grid3 G{m, m * m * m};
  const int grid = mlp_grid(G.n);
  const GridReduce gd = reduce_for_grid(ws_reduce(ws, dot, nullptr), uint64_t(grid));
  hipLaunchKernelGGL(k_mlp_smooth<true>, dim3(grid), dim3(64 * kWaves), 0, h->stream, G, params, wbf, r, dinv, cfeat,
                     omega, jacobi, c, gd);
  finish_reduce(h->stream, gd, uint64_t(grid));
  return check_launch("feat_mlp_direction_f64");
}

int feat_mlp_features_full(feat_handle_t h, uint64_t m, const double* r, const double* dinv, const float* cfeat,
                           uint16_t* X, double* s, double* rho)
{
  if(!h || m < 3 || !r || !dinv || !cfeat || !X || !s || !rho) return fail(FEAT_ERR_INVALID, "feat_mlp_features_full: bad argument");
  Grid3 G{m, m * m * m};
  hipLaunchKernelGGL(k_mlp_features_full, dim3((G.n + 255) / 256), dim3(256), 0, h->stream, G, r, dinv, cfeat, X, s, rho);
  return check_launch("feat_mlp_features_full");
}

int feat_mlp_energy(feat_handle_t h, uint64_t n, int stage, double coef, const float* y, const double* s,
                    const double* rho, const double* e, const double* q, double* ep, float* t)
{
  if(!h || !n || !y || !s || (stage == 0 && (!rho || !ep)) || (stage == 1 && (!q || !t)))
    return fail(FEAT_ERR_INVALID, "feat_mlp_energy: bad argument");
  hipLaunchKernelGGL(k_mlp_energy, dim3((n + 255) / 256), dim3(256), 0, h->stream, n, stage, coef, y, s, rho, e, q, ep, t);
  return check_launch("feat_mlp_energy");
}

int feat_mlp_step_f64(feat_handle_t h, uint64_t n, double* x, double* d, const double* c, const double* q,
                      const double* a, const double* b)
{
  if(!h || !a || !b || (n && (!x || !d || !c || !q))) return fail(FEAT_ERR_INVALID, "feat_mlp_step_f64: null argument");
  if(x == c || d == q || x == d) return fail(FEAT_ERR_ALIAS, "feat_mlp_step_f64: aliasing arguments");
  if(!n) return FEAT_OK;
  const uint64_t g = (n + 4 * 256 - 1) / (4 * 256);
  hipLaunchKernelGGL(k_mlp_step, dim3(g < 8192 ? g : 8192), dim3(256), 0, h->stream, n, x, d, c, q, a, b);
  return check_launch("feat_mlp_step_f64");
}

} // extern "C"
