// Learned culling surrogate: a Fourier-feature MLP per shape family, evaluated with BF16 MFMA
// (gfx950's v_mfma_f32_16x16x32_bf16 over pairs of 16-wide K blocks, an odd last block paired with
// zeros; FP32 accumulation) on wave64.
//
//   u = p * invLen (p: unscaled local point of shape s)
//   x[0:96]   sin/cos(w_j u_a), w_j = pi 2^(j/3), j < 16, index 32a + 2j (+1 for cos),
//             evaluated as sinpi/cospi(u_a 2^(j/3))
//   x[96:112] embedding of s (16)
//   h1..h4 = silu(. W_k + b_k) (112->64, then 64->64 x3), out = h4 . w5 + b5
//   predicted SDF (unscaled units) = out / invLen
//
// Register layout (one pass = 16 samples per wave): lane l holds sample (l % 16) and, for every
// 16-wide feature block kb, features 16 kb + 4 (l / 16) + i, i < 4.  This is both the D layout of
// a transposed product H^T = W^T X^T and the B-operand layout of the next one, so layers chain
// without shuffles.  Weights are pre-swizzled into A-operand order:
//   forward  (W^T, tile (mb = out block, kb = in block)): lane l holds W[16kb + 4(l/16) + i][16mb + l%16]
//   backward (W,   tile (mb = in block,  kb = out block)): lane l holds W[16mb + l%16][16kb + 4(l/16) + i]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsd {

typedef short sur_v4s __attribute__((ext_vector_type(4)));
typedef short sur_v8s __attribute__((ext_vector_type(8)));
typedef __bf16 sur_v8b __attribute__((ext_vector_type(8)));
typedef float sur_v4f __attribute__((ext_vector_type(4)));

// acc + A B over two 16-wide K blocks in one 16x16x32 instruction.  Lane l passes its 4 values of
// block 0 followed by its 4 values of block 1, for A and B alike: the instruction's K index 8 (l / 16)
// + t then names the same K element in both operands, so its 32-term sum is the sum over both
// blocks (in the instruction's own order).
__device__ __forceinline__ sur_v4f surMfma2(sur_v4s a0, sur_v4s a1, sur_v4s b0, sur_v4s b1, sur_v4f acc) {
  const sur_v8s a = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
  const sur_v8s b = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(sur_v8b, a), __builtin_bit_cast(sur_v8b, b), acc,
                                                 0, 0, 0);
}
// acc + sum over K blocks kb < KB of A(tile(kb)) B(kb): pairs of blocks, an odd last block with a
// zero block beside it.  Every instruction of the chain is the same 16x16x32 MFMA: each one takes
// the previous one's result as its accumulator.
template <int KB, class TileA, class OpB>
__device__ __forceinline__ sur_v4f surMfmaK(TileA tileA, OpB opB, sur_v4f acc) {
#pragma unroll
  for (int kb = 0; kb + 1 < KB; kb += 2) acc = surMfma2(tileA(kb), tileA(kb + 1), opB(kb), opB(kb + 1), acc);
  if (KB & 1) {
    const sur_v4s zero = {0, 0, 0, 0};
    acc = surMfma2(tileA(KB - 1), zero, opB(KB - 1), zero, acc);
  }
  return acc;
}

constexpr int kSurFreq = 16;
constexpr int kSurEmb = 16;
constexpr int kSurIn = 6 * kSurFreq + kSurEmb;  // 112
constexpr int kSurHid = 64;
constexpr int kSurKB1 = kSurIn / 16;            // 7 input blocks of layer 1
// flat FP32 parameter vector
constexpr int kSurOffW[4] = {0, 7232, 11392, 15552};
constexpr int kSurOffB[4] = {7168, 11328, 15488, 19648};
constexpr int kSurOffW5 = 19712;
constexpr int kSurOffB5 = 19776;
constexpr int kSurNParam = 19777;
// swizzled BF16 tiles (256 values each): layer 1 has 28 tiles, layers 2..4 16 each
constexpr int kSurTileBase[4] = {0, 28, 44, 60};
constexpr int kSurNTiles = 76;

struct SurNet {
  const sur_v4s* wF;    // forward tiles  [76][64 lanes] of 4 bf16
  const sur_v4s* wB;    // backward tiles [76][64 lanes] of 4 bf16
  const float* param;   // FP32 master parameters (biases, w5, b5 are read from here)
  const uint16_t* emb;  // [nShapes][16] bf16
  float invLen;
  int nShapes;
};

__device__ __forceinline__ uint16_t surBf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // nearest, ties to even (finite inputs)
  return uint16_t(u >> 16);
}
__device__ __forceinline__ float surF32(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ sur_v4s surPack(float a, float b, float c, float d) {
  sur_v4s r;
  r.x = short(surBf16(a));
  r.y = short(surBf16(b));
  r.z = short(surBf16(c));
  r.w = short(surBf16(d));
  return r;
}
__device__ __forceinline__ float surSilu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float surDSilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
// w_j / pi = 2^(j/3): features are sinpi / cospi of u 2^(j/3) (no large-argument reduction).
__device__ __forceinline__ float surFreqOverPi(int j) { return exp2f(float(j) * (1.0f / 3.0f)); }

// Input feature k (< 96) of a sample at normalised position u.
__device__ __forceinline__ float surFeature(int k, float ux, float uy, float uz) {
  const int a = k >> 5, j = (k & 31) >> 1;
  const float u = a == 0 ? ux : (a == 1 ? uy : uz);
  const float t = u * surFreqOverPi(j);
  return (k & 1) ? cospif(t) : sinpif(t);
}

// Input operand blocks of one pass: lane's sample (ux, uy, uz, shape) and its 4 features per block.
__device__ __forceinline__ void surInput(const SurNet& N, float ux, float uy, float uz, int shape, int lane,
                                         sur_v4s x[kSurKB1], float xf[kSurKB1][4]) {
  const int q = lane >> 4;
#pragma unroll
  for (int kb = 0; kb < kSurKB1; ++kb) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = 16 * kb + 4 * q + i;
      xf[kb][i] = kb < 6 ? surFeature(k, ux, uy, uz) : surF32(N.emb[shape * kSurEmb + (k - 96)]);
    }
    x[kb] = surPack(xf[kb][0], xf[kb][1], xf[kb][2], xf[kb][3]);
  }
}

// One dense layer in transposed form: z^T = W^T h^T + b for the pass's 16 samples.
template <int KB>
__device__ __forceinline__ void surLayer(const SurNet& N, int layer, const sur_v4s* in, int lane, float z[4][4]) {
  const int q = lane >> 4;
  const float* b = N.param + kSurOffB[layer];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    const sur_v4f acc = surMfmaK<KB>([&](int kb) { return N.wF[(kSurTileBase[layer] + mb * KB + kb) * 64 + lane]; },
                                     [&](int kb) { return in[kb]; }, sur_v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int i = 0; i < 4; ++i) z[mb][i] = acc[i] + b[16 * mb + 4 * q + i];
  }
}

// Output of the pass for the lane's sample (all 4 lanes of a sample get the same value).
__device__ __forceinline__ float surOutput(const SurNet& N, const float h4[4][4], int lane) {
  const int q = lane >> 4;
  const float* w5 = N.param + kSurOffW5;
  float s = 0.f;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) s += h4[mb][i] * w5[16 * mb + 4 * q + i];
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  return s + N.param[kSurOffB5];
}

// Forward of one pass.  Keeps z of every layer (for training) when Z != nullptr.
__device__ __forceinline__ float surForward(const SurNet& N, float ux, float uy, float uz, int shape, int lane,
                                            sur_v4s x[kSurKB1], float zs[4][4][4], sur_v4s hs[3][4],
                                            float h4[4][4]) {
  float xf[kSurKB1][4];
  surInput(N, ux, uy, uz, shape, lane, x, xf);
  float z[4][4];
  surLayer<kSurKB1>(N, 0, x, lane, z);
#pragma unroll
  for (int layer = 0; layer < 4; ++layer) {
    if (layer > 0) surLayer<4>(N, layer, hs[layer - 1], lane, z);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      float a[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        zs[layer][mb][i] = z[mb][i];
        a[i] = surSilu(z[mb][i]);
      }
      if (layer < 3) hs[layer][mb] = surPack(a[0], a[1], a[2], a[3]);
      else
#pragma unroll
        for (int i = 0; i < 4; ++i) h4[mb][i] = a[i];
    }
  }
  return surOutput(N, h4, lane);
}

// Forward of one pass for inference only (no saved activations): same arithmetic as surForward.
__device__ __forceinline__ float surForwardInfer(const SurNet& N, float ux, float uy, float uz, int shape, int lane) {
  const int q = lane >> 4;
  sur_v4s h[4];
  {
    sur_v4s x[kSurKB1];
#pragma unroll
    for (int kb = 0; kb < kSurKB1; ++kb) {
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = 16 * kb + 4 * q + i;
        v[i] = kb < 6 ? surFeature(k, ux, uy, uz) : surF32(N.emb[shape * kSurEmb + (k - 96)]);
      }
      x[kb] = surPack(v[0], v[1], v[2], v[3]);
    }
    const float* b = N.param + kSurOffB[0];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const sur_v4f acc =
          surMfmaK<kSurKB1>([&](int kb) { return N.wF[(kSurTileBase[0] + mb * kSurKB1 + kb) * 64 + lane]; },
                            [&](int kb) { return x[kb]; }, sur_v4f{0.f, 0.f, 0.f, 0.f});
      h[mb] = surPack(surSilu(acc[0] + b[16 * mb + 4 * q]), surSilu(acc[1] + b[16 * mb + 4 * q + 1]),
                      surSilu(acc[2] + b[16 * mb + 4 * q + 2]), surSilu(acc[3] + b[16 * mb + 4 * q + 3]));
    }
  }
  float s = 0.f;
#pragma unroll 1
  for (int layer = 1; layer < 4; ++layer) {
    const float* b = N.param + kSurOffB[layer];
    const int base = layer == 1 ? kSurTileBase[1] : (layer == 2 ? kSurTileBase[2] : kSurTileBase[3]);
    sur_v4s hn[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const sur_v4f acc = surMfmaK<4>([&](int kb) { return N.wF[(base + mb * 4 + kb) * 64 + lane]; },
                                      [&](int kb) { return h[kb]; }, sur_v4f{0.f, 0.f, 0.f, 0.f});
      float a[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = surSilu(acc[i] + b[16 * mb + 4 * q + i]);
      if (layer < 3) {
        hn[mb] = surPack(a[0], a[1], a[2], a[3]);
      } else {
        const float* w5 = N.param + kSurOffW5;
#pragma unroll
        for (int i = 0; i < 4; ++i) s += a[i] * w5[16 * mb + 4 * q + i];
      }
    }
    if (layer < 3)
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) h[mb] = hn[mb];
  }
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  return s + N.param[kSurOffB5];
}

// Inference for the wave's 64 queries (lane l owns query l): returns the lane's own output.
// ux/uy/uz/shape are the lane's query; lanes with shape < 0 get 0 (their result is ignored).
__device__ __forceinline__ float surInfer64(const SurNet& N, float ux, float uy, float uz, int shape, int lane) {
  float mine = 0.f;
#pragma unroll 1
  for (int pass = 0; pass < 4; ++pass) {
    const int src = 16 * pass + (lane & 15);
    const float sx = __shfl(ux, src, 64), sy = __shfl(uy, src, 64), sz = __shfl(uz, src, 64);
    const int ss = __shfl(shape, src, 64);
    const float o = surForwardInfer(N, sx, sy, sz, ss < 0 ? 0 : ss, lane);
    // lanes 16 pass .. 16 pass + 15 take their sample's result from lane (l & 15)
    const float got = __shfl(o, lane & 15, 64);
    if ((lane >> 4) == pass) mine = got;
  }
  return shape < 0 ? 0.f : mine;
}


// Inference for the lanes of the wave that hold a query (shape >= 0), packed 16 per pass: only
// ceil(queries / 16) passes run.  `idx` is 64 ints of the wave's LDS.  Lanes without a query get 0.
__device__ __forceinline__ float surInferPacked(const SurNet& N, float ux, float uy, float uz, int shape, int lane,
                                                int* idx) {
  const bool q = shape >= 0;
  const unsigned long long m = __ballot(q);
  const int n = __popcll(m);
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int rank = __popcll(m & lt);
  if (q) idx[rank] = lane;
  __builtin_amdgcn_wave_barrier();
  float mine = 0.f;
  for (int pass = 0; pass * 16 < n; ++pass) {
    const int r = 16 * pass + (lane & 15);
    const int src = r < n ? idx[r] : 0;
    const float sx = __shfl(ux, src, 64), sy = __shfl(uy, src, 64), sz = __shfl(uz, src, 64);
    const int ss = __shfl(shape, src, 64);
    const float o = surForwardInfer(N, sx, sy, sz, (r < n && ss >= 0) ? ss : 0, lane);
    const float got = __shfl(o, rank & 15, 64);
    if (q && (rank >> 4) == pass) mine = got;
  }
  __builtin_amdgcn_wave_barrier();
  return mine;
}

}  // namespace lsd
