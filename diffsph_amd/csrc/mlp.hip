// Per-particle MLP of the learned DFSPH warm start on bf16 MFMA (gfx950).
//
// Network: 16 inputs -> 64 -> 64 -> 64 -> 1, ReLU, weights and layer inputs
// rounded to bf16 (round to nearest even), products accumulated in fp32 by
// v_mfma_f32_16x16x16_bf16; biases, pre-activations and the loss in fp32.
//
// One wave processes 16 samples at a time (one MFMA M tile).  For
// v_mfma_f32_16x16x16bf16_1k lane l (c = l%16, q = l/16) holds
//   A[c][4q..4q+3], B[4q..4q+3][c], C[4q..4q+3][c].
// The C layout of a matrix is the A layout of its transpose, so the weight
// gradients dW = H^T dZ take both operands straight from the C layout of
// the forward/backward activations; only the layer inputs of the next
// forward/backward GEMM go through LDS (16x64 bf16 per wave).
#include "common.h"

namespace dsph {

typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kFin = 16;
constexpr int kHid = 64;
// Flat fp32 parameter layout.
constexpr int kOffW1 = 0;                       // [16][64]
constexpr int kOffB1 = kOffW1 + kFin * kHid;    // [64]
constexpr int kOffW2 = kOffB1 + kHid;           // [64][64]
constexpr int kOffB2 = kOffW2 + kHid * kHid;    // [64]
constexpr int kOffW3 = kOffB2 + kHid;           // [64][64]
constexpr int kOffB3 = kOffW3 + kHid * kHid;    // [64]
constexpr int kOffW4 = kOffB3 + kHid;           // [64]
constexpr int kOffB4 = kOffW4 + kHid;           // [1]
constexpr int kNumParams = kOffB4 + 1;          // 9473

// Packed bf16 B operands, one short4 per (tile, lane):
//   W1 (4 n tiles), W2, W3 (16 tiles each, t = kt*4 + nt), W2^T, W3^T.
constexpr int kPkW1 = 0;
constexpr int kPkW2 = kPkW1 + 4 * 64;
constexpr int kPkW3 = kPkW2 + 16 * 64;
constexpr int kPkW2T = kPkW3 + 16 * 64;
constexpr int kPkW3T = kPkW2T + 16 * 64;
constexpr int kPkTotal = kPkW3T + 16 * 64;      // short4 entries

// float -> bf16 bits, round to nearest even (NaN stays NaN).
__device__ __forceinline__ short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((unsigned int)(unsigned short)b) << 16); }

__device__ __forceinline__ float4v mfma16(short4v a, short4v b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Pack the fp32 master weights into the per-lane B-operand layout.
__global__ void __launch_bounds__(kBlock) k_mlp_pack(const float* __restrict__ prm, short4v* __restrict__ pk) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kPkTotal) return;
  const int l = t & 63, c = l & 15, q = l >> 4;
  const int tile = t >> 6;
  short4v v;
  if (t < kPkW2) {  // W1: tile = nt, K = 16 rows
    const int nt = tile;
    for (int j = 0; j < 4; ++j) v[j] = f2bf(prm[kOffW1 + (4 * q + j) * kHid + 16 * nt + c]);
  } else {
    int base, off, tt;
    bool trans;
    if (t < kPkW3) {
      base = kPkW2; off = kOffW2; trans = false;
    } else if (t < kPkW2T) {
      base = kPkW3; off = kOffW3; trans = false;
    } else if (t < kPkW3T) {
      base = kPkW2T; off = kOffW2; trans = true;
    } else {
      base = kPkW3T; off = kOffW3; trans = true;
    }
    tt = (t - base) >> 6;
    const int kt = tt >> 2, nt = tt & 3;
    for (int j = 0; j < 4; ++j) {
      const int kk = 16 * kt + 4 * q + j, nn = 16 * nt + c;
      // B[kk][nn] = W[kk][nn] (forward) or W^T[kk][nn] = W[nn][kk] (backward)
      v[j] = f2bf(trans ? prm[off + nn * kHid + kk] : prm[off + kk * kHid + nn]);
    }
  }
  pk[t] = v;
}

struct MlpTile {
  // C-layout fp32 pre-activations (z + b) and bf16 post-activations of the
  // three hidden layers: [layer][nt] -> 4 rows (4q + j) of column 16nt + c.
  float4v z[3][4];
  short4v h[3][4];
  float y[4];
};

// LDS transpose: C layout (rows 4q+j, col 16nt+c) -> A layout (row c, cols
// 16kt+4q+j).  buf holds 16 x 64 bf16 for this wave (row stride 68 to keep
// the 16 rows of a column in distinct banks).
constexpr int kLdsStride = 68;

__device__ __forceinline__ void c_to_a(const short4v (&in)[4], short4v (&out)[4], short* buf) {
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int j = 0; j < 4; ++j) buf[(4 * q + j) * kLdsStride + 16 * nt + c] = in[nt][j];
  __syncthreads();
#pragma unroll
  for (int kt = 0; kt < 4; ++kt)
#pragma unroll
    for (int j = 0; j < 4; ++j) out[kt][j] = buf[c * kLdsStride + 16 * kt + 4 * q + j];
  __syncthreads();
}

// Forward of one 16-sample tile starting at row m0 (rows >= n read as 0).
__device__ __forceinline__ void mlp_forward_tile(const float* __restrict__ X, int n, int m0,
                                                 const float* __restrict__ prm, const short4v* __restrict__ pk,
                                                 short* lds, MlpTile& T) {
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  short4v xa;
  {
    const int m = m0 + c;
#pragma unroll
    for (int j = 0; j < 4; ++j) xa[j] = f2bf(m < n ? X[(int64_t)m * kFin + 4 * q + j] : 0.0f);
  }
  // layer 1
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    float4v acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma16(xa, pk[kPkW1 + nt * 64 + l], acc);
    const float b = prm[kOffB1 + 16 * nt + c];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float zz = acc[j] + b;
      T.z[0][nt][j] = zz;
      T.h[0][nt][j] = f2bf(zz > 0.0f ? zz : 0.0f);
    }
  }
  // layers 2, 3
#pragma unroll
  for (int L = 1; L < 3; ++L) {
    short4v a[4];
    c_to_a(T.h[L - 1], a, lds);
    const int pkb = L == 1 ? kPkW2 : kPkW3;
    const int bo = L == 1 ? kOffB2 : kOffB3;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      float4v acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) acc = mfma16(a[kt], pk[pkb + (kt * 4 + nt) * 64 + l], acc);
      const float b = prm[bo + 16 * nt + c];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float zz = acc[j] + b;
        T.z[L][nt][j] = zz;
        T.h[L][nt][j] = f2bf(zz > 0.0f ? zz : 0.0f);
      }
    }
  }
  // output layer (VALU): y[4q+j] = sum_k h3[4q+j][k] w4[k] + b4
  float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float w = bf2f(f2bf(prm[kOffW4 + 16 * nt + c]));
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] += bf2f(T.h[2][nt][j]) * w;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float v = s[j];
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    T.y[j] = v + prm[kOffB4];
  }
}

// Inference on the transposed product, Z^T = W^T H^T: hidden units are the
// M rows and the 16 samples of a wave tile the N columns.  The C layout of a
// layer's output (lane (c, q) holds rows 4q..4q+3 of an M tile for sample c)
// is then already the B layout of the next layer's K chunk, so activations
// never leave registers: no LDS and no barriers.  Layer 1 (K = 16 inputs)
// runs on v_mfma_f32_16x16x16_bf16; layers 2 and 3 (K = 64) on gfx950's
// v_mfma_f32_16x16x32_bf16, two K chunks each.  A K chunk of 32 hidden units
// takes rows 4q..4q+3 of two consecutive M tiles into lane group q's eight
// K slots; the weight (A) operand uses the same permutation of K, so the
// products pair up (the sum over K does not depend on its order in the
// operands beyond fp32 rounding).  The weight fragments are the packed B
// operands of the training kernels read as A operands (the two layouts are
// mirror images): W1 tile nt is lane (c, q)'s W1[4q+j][16nt+c] =
// W1^T[16nt+c][4q+j]; W2/W3 tiles (kt, nt) give the halves of a K chunk.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ bf16x8 cat8(bf16x4 lo, bf16x4 hi) {
  return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}
__device__ __forceinline__ bf16x8 cat8(short4v lo, short4v hi) {
  return cat8(__builtin_bit_cast(bf16x4, lo), __builtin_bit_cast(bf16x4, hi));
}

// relu(z) rounded to bf16 (round to nearest even: the hardware conversion,
// v_cvt_pk_bf16_f32, two values per instruction).  The bias is already in
// z: the MFMA accumulator starts from it.  max(z, 0) is taken on the bit
// pattern as a signed integer (one instruction; negative values and -0
// become +0, positive values are kept).
__device__ __forceinline__ float relu_bits(float z) {
  return __int_as_float(max(__float_as_int(z), 0));
}
__device__ __forceinline__ bf16x4 relu_bf16(float4v acc) {
  float2v lo, hi;
  lo[0] = relu_bits(acc[0]);
  lo[1] = relu_bits(acc[1]);
  hi[0] = relu_bits(acc[2]);
  hi[1] = relu_bits(acc[3]);
  const bf16x2 l2 = __builtin_convertvector(lo, bf16x2), h2 = __builtin_convertvector(hi, bf16x2);
  return __builtin_shufflevector(l2, h2, 0, 1, 2, 3);
}

__device__ __forceinline__ float4v mfma32(bf16x8 a, bf16x8 b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct InferW {
  short4v w1[4];
  bf16x8 w2[4][2], w3[4][2];
  bf16x8 w4[2];  // output row: w4 in row 0 of the A tile, zeros elsewhere
  float4v b1[4], b2[4], b3[4];
  float4v b4;
};

// One 16-sample tile from its inputs (lane (c, q): sample c, inputs
// 4q..4q+3).  Every layer, the output one included, is an MFMA whose
// accumulator starts from the bias; returns the output tile, whose row 0
// (lane (c, 0), element 0) is y of sample c.
__device__ __forceinline__ float4v infer_tile(const InferW& W, float4 xv) {
  float2v x01, x23;
  x01[0] = xv.x;
  x01[1] = xv.y;
  x23[0] = xv.z;
  x23[1] = xv.w;
  const bf16x4 xb = __builtin_shufflevector(__builtin_convertvector(x01, bf16x2),
                                            __builtin_convertvector(x23, bf16x2), 0, 1, 2, 3);
  const short4v xs = __builtin_bit_cast(short4v, xb);
  bf16x4 h[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) h[mt] = relu_bf16(mfma16(W.w1[mt], xs, W.b1[mt]));
#pragma unroll
  for (int L = 0; L < 2; ++L) {
    const bf16x8 k0 = cat8(h[0], h[1]), k1 = cat8(h[2], h[3]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      float4v acc = L == 0 ? W.b2[mt] : W.b3[mt];
      acc = mfma32(L == 0 ? W.w2[mt][0] : W.w3[mt][0], k0, acc);
      acc = mfma32(L == 0 ? W.w2[mt][1] : W.w3[mt][1], k1, acc);
      h[mt] = relu_bf16(acc);
    }
  }
  float4v y = mfma32(W.w4[0], cat8(h[0], h[1]), W.b4);
  return mfma32(W.w4[1], cat8(h[2], h[3]), y);
}

// Grid-stride over 16-sample wave tiles, two tiles per pass so that their
// independent MFMA chains overlap, the next pass's inputs in flight (no
// block-wide synchronisation).
// Output epilogue (base != NULL): Y = base + (y * scale + shift), negative
// values set to 0 when clamp (NaN kept).
struct FwdOut {
  const float* base;
  float scale, shift;
  int clamp;
  __device__ __forceinline__ float operator()(float y, int m) const {
    if (!base) return y;
    const float v = base[m] + (y * scale + shift);
    return (clamp && v < 0.0f) ? 0.0f : v;
  }
};

__global__ void __launch_bounds__(kBlock) k_mlp_forward(const float* __restrict__ X, int n,
                                                        const float* __restrict__ prm,
                                                        const short4v* __restrict__ pk, float* __restrict__ Y,
                                                        FwdOut out) {
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  InferW W;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    W.w1[mt] = pk[kPkW1 + mt * 64 + l];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
      W.w2[mt][kc] = cat8(pk[kPkW2 + ((2 * kc) * 4 + mt) * 64 + l], pk[kPkW2 + ((2 * kc + 1) * 4 + mt) * 64 + l]);
      W.w3[mt][kc] = cat8(pk[kPkW3 + ((2 * kc) * 4 + mt) * 64 + l], pk[kPkW3 + ((2 * kc + 1) * 4 + mt) * 64 + l]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 16 * mt + 4 * q + j;
      W.b1[mt][j] = prm[kOffB1 + r];
      W.b2[mt][j] = prm[kOffB2 + r];
      W.b3[mt][j] = prm[kOffB3 + r];
    }
  }
  // output A tile: row 0 (lanes c == 0) holds w4 at the K chunk's hidden
  // units (32 kc + 4q + j, then 32 kc + 16 + 4q + j), the other rows zero
#pragma unroll
  for (int kc = 0; kc < 2; ++kc) {
    short v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int r = 32 * kc + (j < 4 ? 4 * q + j : 16 + 4 * q + (j - 4));
      v[j] = c == 0 ? f2bf(prm[kOffW4 + r]) : (short)0;
    }
    W.w4[kc] = __builtin_bit_cast(bf16x8, v);
  }
  W.b4 = float4v{prm[kOffB4], 0.f, 0.f, 0.f};
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int ntiles = (n + 15) / 16;
  auto load = [&](int t) {
    const int m = t * 16 + c;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t < ntiles && m < n) v = *reinterpret_cast<const float4*>(X + (int64_t)m * kFin + 4 * q);
    return v;
  };
  // the next pass's inputs are loaded while this pass computes
  float4 n0 = load(wave), n1 = load(wave + nwaves);
  for (int t = wave; t < ntiles; t += 2 * nwaves) {
    const int t1 = t + nwaves;
    const float4 x0 = n0, x1 = n1;
    n0 = load(t + 2 * nwaves);
    n1 = load(t1 + 2 * nwaves);
    const float4v y0 = infer_tile(W, x0);
    const float4v y1 = infer_tile(W, x1);
    const int m0 = t * 16 + c, m1 = t1 * 16 + c;
    if (q == 0 && m0 < n) Y[m0] = out(y0[0], m0);
    if (q == 0 && t1 < ntiles && m1 < n) Y[m1] = out(y1[0], m1);
  }
}

// Training: forward + backward of the loss sum_m lw (y_m - t_m)^2 over all
// samples; per-block partial gradients (fixed order, deterministic) and
// partial loss.  gpart: [gridDim.x][kNumParams], lpart: [gridDim.x].
__global__ void __launch_bounds__(kBlock) k_mlp_train(const float* __restrict__ X, const float* __restrict__ Tg,
                                                      int n, float lw, const float* __restrict__ prm,
                                                      const short4v* __restrict__ pk, float* __restrict__ gpart,
                                                      float* __restrict__ lpart) {
  __shared__ short lds[4][16 * kLdsStride];
  __shared__ float red[4][1024];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  float4v gW1[4], gW2[4][4], gW3[4][4];
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3[4] = {0, 0, 0, 0}, gw4[4] = {0, 0, 0, 0};
  float gb4 = 0.f, loss = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gW1[a] = float4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      gW2[a][b] = float4v{0.f, 0.f, 0.f, 0.f};
      gW3[a][b] = float4v{0.f, 0.f, 0.f, 0.f};
    }
  }
  float w4b[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) w4b[nt] = bf2f(f2bf(prm[kOffW4 + 16 * nt + c]));

  const int ntiles = (n + 63) / 64;
  for (int bt = blockIdx.x; bt < ntiles; bt += gridDim.x) {
    const int m0 = bt * 64 + w * 16;
    MlpTile T;
    mlp_forward_tile(X, n, m0, prm, pk, lds[w], T);
    // dL/dy
    float dy[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + 4 * q + j;
      const float e = m < n ? T.y[j] - Tg[m] : 0.0f;
      dy[j] = m < n ? 2.0f * lw * e : 0.0f;
      if (c == 0) loss += lw * e * e;
      if (c == 0) gb4 += dy[j];
    }
    // output layer: dw4 and dz3
    short4v dzb[4];
    float4v dz[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        gw4[nt] += dy[j] * bf2f(T.h[2][nt][j]);
        const float d = T.z[2][nt][j] > 0.0f ? dy[j] * w4b[nt] : 0.0f;
        dz[nt][j] = d;
        dzb[nt][j] = f2bf(d);
        gb3[nt] += d;
      }
    }
    // layers 3, 2, 1
#pragma unroll
    for (int L = 2; L >= 1; --L) {
      // dW_{L+1} += H_L^T dZ_{L+1} (both operands in C layout)
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
          if (L == 2)
            gW3[kt][nt] = mfma16(T.h[1][kt], dzb[nt], gW3[kt][nt]);
          else
            gW2[kt][nt] = mfma16(T.h[0][kt], dzb[nt], gW2[kt][nt]);
        }
      // dH_L = dZ_{L+1} W_{L+1}^T
      short4v a[4];
      c_to_a(dzb, a, lds[w]);
      const int pkb = L == 2 ? kPkW3T : kPkW2T;
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        float4v acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) acc = mfma16(a[kt], pk[pkb + (kt * 4 + nt) * 64 + l], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float d = T.z[L - 1][nt][j] > 0.0f ? acc[j] : 0.0f;
          dz[nt][j] = d;
          dzb[nt][j] = f2bf(d);
          if (L == 2)
            gb2[nt] += d;
          else
            gb1[nt] += d;
        }
      }
    }
    // dW1 += X^T dZ1: X in C layout (rows 4q+j, column c)
    short4v xc;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + 4 * q + j;
      xc[j] = f2bf(m < n ? X[(int64_t)m * kFin + c] : 0.0f);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) gW1[nt] = mfma16(xc, dzb[nt], gW1[nt]);
  }

  // Reduce the 4 waves of the block in fixed order through LDS, one
  // parameter group at a time, and write the block partial.
  float* out = gpart + (int64_t)blockIdx.x * kNumParams;
  float* r = red[w];
  // Helper lambdas write this wave's contribution of a [rows x 64] matrix
  // chunk into r (indexed by row*64 + col), then wave 0 sums.
  auto reduce_chunk = [&](int count, int dst) {
    __syncthreads();
    if (w == 0) {
      for (int k = l; k < count; k += 64) {
        float v = red[0][k] + red[1][k];
        v = v + red[2][k];
        v = v + red[3][k];
        out[dst + k] = v;
      }
    }
    __syncthreads();
  };
  // W1 (16 x 64)
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int j = 0; j < 4; ++j) r[(4 * q + j) * kHid + 16 * nt + c] = gW1[nt][j];
  reduce_chunk(kFin * kHid, kOffW1);
  // W2, W3 in 4 chunks of 16 rows each
#pragma unroll
  for (int mat = 0; mat < 2; ++mat) {
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) r[(4 * q + j) * kHid + 16 * nt + c] = mat == 0 ? gW2[kt][nt][j] : gW3[kt][nt][j];
      reduce_chunk(16 * kHid, (mat == 0 ? kOffW2 : kOffW3) + 16 * kt * kHid);
    }
  }
  // biases, w4, b4, loss: sum the 4 quarter-lanes first
  float vb1[4], vb2[4], vb3[4], vw4[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    vb1[nt] = gb1[nt];
    vb2[nt] = gb2[nt];
    vb3[nt] = gb3[nt];
    vw4[nt] = gw4[nt];
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      vb1[nt] += __shfl_xor(vb1[nt], o, 64);
      vb2[nt] += __shfl_xor(vb2[nt], o, 64);
      vb3[nt] += __shfl_xor(vb3[nt], o, 64);
      vw4[nt] += __shfl_xor(vw4[nt], o, 64);
    }
  }
  float vb4 = gb4, vl = loss;
#pragma unroll
  for (int o = 16; o < 64; o <<= 1) {
    vb4 += __shfl_xor(vb4, o, 64);
    vl += __shfl_xor(vl, o, 64);
  }
  if (q == 0) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      r[16 * nt + c] = vb1[nt];
      r[64 + 16 * nt + c] = vb2[nt];
      r[128 + 16 * nt + c] = vb3[nt];
      r[192 + 16 * nt + c] = vw4[nt];
    }
    if (c == 0) {
      r[256] = vb4;
      r[257] = vl;
    }
  }
  __syncthreads();
  if (w == 0) {
    for (int k = l; k < 258; k += 64) {
      float v = red[0][k] + red[1][k];
      v = v + red[2][k];
      v = v + red[3][k];
      if (k < 64)
        out[kOffB1 + k] = v;
      else if (k < 128)
        out[kOffB2 + k - 64] = v;
      else if (k < 192)
        out[kOffB3 + k - 128] = v;
      else if (k < 256)
        out[kOffW4 + k - 192] = v;
      else if (k == 256)
        out[kOffB4] = v;
      else
        lpart[blockIdx.x] = v;
    }
  }
}

// Sum block partials in a fixed two-level order: group g of kRedGroups sums
// its contiguous run of blocks, then the group sums are added in order.
// Column kNumParams of a partial row is the block's loss.
constexpr int kRedGroups = 32;

__global__ void __launch_bounds__(kBlock) k_mlp_reduce_groups(const float* __restrict__ gpart,
                                                              const float* __restrict__ lpart, int nb,
                                                              float* __restrict__ gmid) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  if (p > kNumParams) return;
  const int chunk = (nb + kRedGroups - 1) / kRedGroups;
  const int b0 = g * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
  float s = 0.f;
  if (p < kNumParams) {
    for (int b = b0; b < b1; ++b) s += gpart[(int64_t)b * kNumParams + p];
  } else {
    for (int b = b0; b < b1; ++b) s += lpart[b];
  }
  gmid[(int64_t)g * (kNumParams + 1) + p] = s;
}

__global__ void __launch_bounds__(kBlock) k_mlp_reduce_final(const float* __restrict__ gmid,
                                                             float* __restrict__ grad, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kNumParams) return;
  float s = 0.f;
  for (int g = 0; g < kRedGroups; ++g) s += gmid[(int64_t)g * (kNumParams + 1) + p];
  if (p < kNumParams)
    grad[p] = s;
  else if (loss)
    loss[0] = s;
}

// Adam (fp32 master weights), bias-corrected.
__global__ void __launch_bounds__(kBlock) k_adam(float* __restrict__ prm, const float* __restrict__ g,
                                                 float* __restrict__ m1, float* __restrict__ m2, int n, float lr,
                                                 float b1, float b2, float eps, float c1, float c2, float wd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float gg = g[p] + wd * prm[p];
  float a = b1 * m1[p] + (1.0f - b1) * gg;
  float b = b2 * m2[p] + (1.0f - b2) * gg * gg;
  m1[p] = a;
  m2[p] = b;
  prm[p] = prm[p] - lr * (a / c1) / (sqrtf(b / c2) + eps);
}

}  // namespace dsph

using namespace dsph;

extern "C" int dsph_mlp_num_params(void) { return kNumParams; }
extern "C" int dsph_mlp_num_inputs(void) { return kFin; }
extern "C" size_t dsph_mlp_packed_bytes(void) { return (size_t)kPkTotal * sizeof(short4v); }

extern "C" int dsph_mlp_pack(const float* params, void* packed, void* stream) {
  if (!params || !packed) return DSPH_ERR_ARG;
  k_mlp_pack<<<grid_for(kPkTotal), kBlock, 0, as_stream(stream)>>>(params, static_cast<short4v*>(packed));
  return launch_status();
}

static int mlp_grid(int64_t n) {
  int64_t tiles = (n + 63) / 64;
  int64_t g = tiles < 1024 ? tiles : 1024;
  return (int)(g < 1 ? 1 : g);
}

// Inference grid: four 16-sample wave tiles per 256-thread block pass, at
// most 512 blocks (two per CU at the kernel's occupancy of two waves per
// SIMD): every wave loads the ~40 KB of weight fragments once and then
// strides over its tiles, so the weight reads stay small next to the input.
static int mlp_forward_grid(int64_t n) {
  const int64_t tiles = (n + 15) / 16;
  int64_t g = (tiles + 3) / 4;
  if (g < 1) g = 1;
  if (g > 512) g = 512;
  return (int)g;
}

extern "C" int dsph_mlp_forward(const float* X, int64_t n, const float* params, const void* packed, float* Y,
                                void* stream) {
  if (n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  if (!X || !params || !packed || !Y) return DSPH_ERR_ARG;
  if (reinterpret_cast<uintptr_t>(X) % 16 != 0) return DSPH_ERR_ARG;  // rows are read as float4
  k_mlp_forward<<<mlp_forward_grid(n), kBlock, 0, as_stream(stream)>>>(
      X, (int)n, params, static_cast<const short4v*>(packed), Y, FwdOut{nullptr, 1.0f, 0.0f, 0});
  return launch_status();
}

extern "C" int dsph_mlp_forward_affine(const float* X, int64_t n, const float* params, const void* packed,
                                       const float* base, float scale, float shift, int32_t clamp, float* Y,
                                       void* stream) {
  if (n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  if (!X || !params || !packed || !Y || !base) return DSPH_ERR_ARG;
  if (reinterpret_cast<uintptr_t>(X) % 16 != 0) return DSPH_ERR_ARG;
  k_mlp_forward<<<mlp_forward_grid(n), kBlock, 0, as_stream(stream)>>>(
      X, (int)n, params, static_cast<const short4v*>(packed), Y, FwdOut{base, scale, shift, clamp ? 1 : 0});
  return launch_status();
}

extern "C" size_t dsph_mlp_train_workspace(int64_t n) {
  return (((size_t)mlp_grid(n) + kRedGroups) * (kNumParams + 1) * sizeof(float) + 255) / 256 * 256;
}

extern "C" int dsph_mlp_loss_grad(const float* X, const float* target, int64_t n, float loss_weight,
                                  const float* params, const void* packed, float* grad, float* loss, void* ws,
                                  size_t ws_bytes, void* stream) {
  if (n < 0 || n >= INT32_MAX || !grad) return DSPH_ERR_ARG;
  if (n > 0 && (!X || !target || !params || !packed)) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_mlp_train_workspace(n)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  const int nb = mlp_grid(n);
  float* gpart = static_cast<float*>(ws);
  float* lpart = gpart + (size_t)nb * kNumParams;
  if (n == 0) {
    DSPH_CHECK_HIP(hipMemsetAsync(grad, 0, sizeof(float) * kNumParams, st));
    if (loss) DSPH_CHECK_HIP(hipMemsetAsync(loss, 0, sizeof(float), st));
    return DSPH_OK;
  }
  k_mlp_train<<<nb, kBlock, 0, st>>>(X, target, (int)n, loss_weight, params, static_cast<const short4v*>(packed),
                                     gpart, lpart);
  float* gmid = lpart + nb;
  k_mlp_reduce_groups<<<dim3(grid_for(kNumParams + 1), kRedGroups), kBlock, 0, st>>>(gpart, lpart, nb, gmid);
  k_mlp_reduce_final<<<grid_for(kNumParams + 1), kBlock, 0, st>>>(gmid, grad, loss);
  return launch_status();
}

extern "C" int dsph_adam_step(float* params, const float* grad, float* m1, float* m2, int64_t n, float lr,
                              float beta1, float beta2, float eps, float weight_decay, int32_t step, void* stream) {
  if (n < 0 || n >= INT32_MAX || step < 1) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  if (!params || !grad || !m1 || !m2) return DSPH_ERR_ARG;
  const float c1 = (float)(1.0 - std::pow((double)beta1, (double)step));
  const float c2 = (float)(1.0 - std::pow((double)beta2, (double)step));
  k_adam<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(params, grad, m1, m2, (int)n, lr, beta1, beta2, eps, c1, c2,
                                                        weight_decay);
  return launch_status();
}
