// Training and evaluation kernels of the learned culling surrogate (network in surrogate.h).
//
//   k_sur_sample   training / validation samples from the shapes' own SDF grids: grid nodes
//                  (exact values) and uniform points of the grid box (FP32 trilinear values),
//                  targets clamped to [-clamp, clamp] in normalised units
//   k_sur_train    one optimizer step's forward + backward over 64 samples per block; per-block
//                  partial gradients (deterministic, no atomics)
//   k_sur_adam     sums the block partials in a fixed order, Adam update of the FP32 masters,
//                  refreshes the BF16 swizzled copies
//   k_sur_eval     outputs and |error| statistics of a sample set
#include <hip/hip_runtime.h>

#include "surrogate_train.h"

namespace lsd {



__device__ __forceinline__ unsigned long long surHash(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ float surUnit(unsigned long long h) { return float(h >> 40) * (1.0f / 16777216.0f); }

// FP32 trilinear value of a grid at index coordinates g (inside the box).
__device__ __forceinline__ float surTri(const float* g0, int nx, int ny, float gx, float gy, float gz, int mx, int my,
                                        int mz) {
  const int i0 = min(max(int(floorf(gx)), 0), mx - 2), j0 = min(max(int(floorf(gy)), 0), my - 2),
            k0 = min(max(int(floorf(gz)), 0), mz - 2);
  const float x = fminf(fmaxf(gx - i0, 0.f), 1.f), y = fminf(fmaxf(gy - j0, 0.f), 1.f), z = fminf(fmaxf(gz - k0, 0.f), 1.f);
  const int sy = nx, sz = nx * ny;
  const float* g = g0 + i0 + j0 * sy + k0 * sz;
  const float a = (g[0] * (1.f - x) + g[1] * x) * (1.f - y) + (g[sy] * (1.f - x) + g[sy + 1] * x) * y;
  const float b = (g[sz] * (1.f - x) + g[sz + 1] * x) * (1.f - y) + (g[sz + sy] * (1.f - x) + g[sz + sy + 1] * x) * y;
  return a * (1.f - z) + b * z;
}

__global__ void k_sur_sample(SurFamily F, SurSamples S, unsigned long long seed) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < S.n;
       i += (long long)gridDim.x * blockDim.x) {
    const int slot = int(i % F.nShapes);
    const ShapeDev sh = F.shapes[F.famShape[slot]];
    const float* g0 = F.sdf + sh.sdfOff;
    const unsigned long long h0 = surHash(seed ^ (unsigned long long)i), h1 = surHash(h0), h2 = surHash(h1);
    float gx, gy, gz, phi;
    if (((i / F.nShapes) & 1) == 0) {  // a grid node, exact value
      const int ix = int(h0 % (unsigned long long)sh.nx), iy = int(h1 % (unsigned long long)sh.ny),
                iz = int(h2 % (unsigned long long)sh.nz);
      gx = float(ix); gy = float(iy); gz = float(iz);
      phi = g0[ix + sh.nx * (iy + sh.ny * iz)];
    } else {  // uniform in the grid box
      gx = surUnit(h0) * float(sh.nx - 1);
      gy = surUnit(h1) * float(sh.ny - 1);
      gz = surUnit(h2) * float(sh.nz - 1);
      phi = surTri(g0, sh.nx, sh.ny, gx, gy, gz, sh.nx, sh.ny, sh.nz);
    }
    S.x[3 * i + 0] = float(sh.ox + double(gx) * sh.h);
    S.x[3 * i + 1] = float(sh.oy + double(gy) * sh.h);
    S.x[3 * i + 2] = float(sh.oz + double(gz) * sh.h);
    S.slot[i] = slot;
    S.t[i] = fminf(fmaxf(phi * F.invLen, -F.clamp), F.clamp);
  }
}



__global__ __launch_bounds__(256) void k_sur_train(SurTrainArgs A) {
  __shared__ uint16_t sX[kSurRows][kSurIn];
  __shared__ uint16_t sH[3][kSurRows][kSurHid];
  __shared__ uint16_t sDZ[4][kSurRows][kSurHid];
  __shared__ float sRed[4][4 * kSurHid + kSurHid + 2];  // per wave: db1..db4, dw5, db5, loss
  __shared__ float sEmb[kSurRows][kSurEmb];
  __shared__ int sSlot[kSurRows];
  const SurNet& N = A.net;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane >> 4;
  const int row = 16 * w + (lane & 15);
  const long long r = (long long)blockIdx.x * kSurRows + row;
  const long long si = (A.data.n > 0) ? (A.start + r * A.stride) % A.data.n : 0;
  const float px = A.data.x[3 * si], py = A.data.x[3 * si + 1], pz = A.data.x[3 * si + 2];
  const int slot = A.data.slot[si];
  const float target = A.data.t[si];
  // ---- forward ----
  sur_v4s x[kSurKB1];
  float zs[4][4][4];
  sur_v4s hs[3][4];
  float h4[4][4];
  const float out = surForward(N, px * N.invLen, py * N.invLen, pz * N.invLen, slot, lane, x, zs, hs, h4);
#pragma unroll
  for (int kb = 0; kb < kSurKB1; ++kb) {
    sX[row][16 * kb + 4 * q + 0] = uint16_t(x[kb].x);
    sX[row][16 * kb + 4 * q + 1] = uint16_t(x[kb].y);
    sX[row][16 * kb + 4 * q + 2] = uint16_t(x[kb].z);
    sX[row][16 * kb + 4 * q + 3] = uint16_t(x[kb].w);
  }
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      sH[l][row][16 * mb + 4 * q + 0] = uint16_t(hs[l][mb].x);
      sH[l][row][16 * mb + 4 * q + 1] = uint16_t(hs[l][mb].y);
      sH[l][row][16 * mb + 4 * q + 2] = uint16_t(hs[l][mb].z);
      sH[l][row][16 * mb + 4 * q + 3] = uint16_t(hs[l][mb].w);
    }
  if (q == 0) sSlot[row] = slot;
  // ---- backward through the wave's own 16 samples ----
  const float err = out - target;
  const float dout = 2.0f * err / float(A.batch);
  float dz[4][4];
  const float* w5 = N.param + kSurOffW5;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) dz[mb][i] = dout * w5[16 * mb + 4 * q + i] * surDSilu(zs[3][mb][i]);
  // sums over the wave's 16 samples (lanes with equal q): xor over the low 4 lane bits
  auto sum16 = [&](float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
  };
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float g5 = sum16(h4[mb][i] * dout);
      if ((lane & 15) == 0) sRed[w][4 * kSurHid + 16 * mb + 4 * q + i] = g5;
    }
  {
    const float g = sum16(dout);
    const float l2 = sum16(err * err);
    if (lane == 0) {
      sRed[w][5 * kSurHid] = g;
      sRed[w][5 * kSurHid + 1] = l2 / float(A.batch);
    }
  }
#pragma unroll
  for (int layer = 3; layer >= 0; --layer) {
    sur_v4s dzp[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float gb = sum16(dz[mb][i]);
        if ((lane & 15) == 0) sRed[w][layer * kSurHid + 16 * mb + 4 * q + i] = gb;
      }
      dzp[mb] = surPack(dz[mb][0], dz[mb][1], dz[mb][2], dz[mb][3]);
      sDZ[layer][row][16 * mb + 4 * q + 0] = uint16_t(dzp[mb].x);
      sDZ[layer][row][16 * mb + 4 * q + 1] = uint16_t(dzp[mb].y);
      sDZ[layer][row][16 * mb + 4 * q + 2] = uint16_t(dzp[mb].z);
      sDZ[layer][row][16 * mb + 4 * q + 3] = uint16_t(dzp[mb].w);
    }
    if (layer > 0) {
      // dh_prev^T = W dz^T, then through the previous SiLU
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const sur_v4f acc =
            surMfmaK<4>([&](int kb) { return N.wB[(kSurTileBase[layer] + mb * 4 + kb) * 64 + lane]; },
                        [&](int kb) { return dzp[kb]; }, sur_v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
        for (int i = 0; i < 4; ++i) dz[mb][i] = acc[i] * surDSilu(zs[layer - 1][mb][i]);
      }
    } else {
      // embedding rows 96..111 of layer 1: input block 6
      const sur_v4f acc = surMfmaK<4>([&](int kb) { return N.wB[(kSurTileBase[0] + 6 * 4 + kb) * 64 + lane]; },
                                      [&](int kb) { return dzp[kb]; }, sur_v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int i = 0; i < 4; ++i) sEmb[row][4 * q + i] = acc[i];
    }
  }
  __syncthreads();
  // ---- weight gradients over the block's 64 samples: dW = H^T dZ (samples are the K dimension) ----
  float* G = A.partial + (long long)blockIdx.x * (kSurNParam + N.nShapes * kSurEmb + 1);
  const int m16 = lane & 15;
  for (int t = w; t < kSurNTiles; t += 4) {
    const int layer = t < 28 ? 0 : (t < 44 ? 1 : (t < 60 ? 2 : 3));
    const int local = t - kSurTileBase[layer];
    const int kb = local >> 2, nb = local & 3;
    // (the K dimension is the block's 64 samples, in 4 blocks of 16)
    const int k = 16 * kb + m16, n = 16 * nb + m16;
    auto opA = [&](int p) {
      const int s0 = 16 * p + 4 * q;
      sur_v4s a;
      if (layer == 0) {
        a.x = short(sX[s0][k]); a.y = short(sX[s0 + 1][k]); a.z = short(sX[s0 + 2][k]); a.w = short(sX[s0 + 3][k]);
      } else {
        a.x = short(sH[layer - 1][s0][k]); a.y = short(sH[layer - 1][s0 + 1][k]);
        a.z = short(sH[layer - 1][s0 + 2][k]); a.w = short(sH[layer - 1][s0 + 3][k]);
      }
      return a;
    };
    auto opB = [&](int p) {
      const int s0 = 16 * p + 4 * q;
      sur_v4s b;
      b.x = short(sDZ[layer][s0][n]); b.y = short(sDZ[layer][s0 + 1][n]);
      b.z = short(sDZ[layer][s0 + 2][n]); b.w = short(sDZ[layer][s0 + 3][n]);
      return b;
    };
    const sur_v4f acc = surMfmaK<4>(opA, opB, sur_v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int i = 0; i < 4; ++i) G[kSurOffW[layer] + (16 * kb + 4 * q + i) * kSurHid + 16 * nb + m16] = acc[i];
  }
  // biases, w5, b5, loss: the 4 waves' partials in a fixed order
  for (int k = threadIdx.x; k < 5 * kSurHid + 2; k += blockDim.x) {
    const float v = sRed[0][k] + sRed[1][k] + sRed[2][k] + sRed[3][k];
    if (k < 4 * kSurHid) G[kSurOffB[k / kSurHid] + (k % kSurHid)] = v;
    else if (k < 5 * kSurHid) G[kSurOffW5 + (k - 4 * kSurHid)] = v;
    else if (k == 5 * kSurHid) G[kSurOffB5] = v;
    else G[kSurNParam + N.nShapes * kSurEmb] = v;
  }
  // embedding gradients per family slot, samples in row order
  for (int k = threadIdx.x; k < N.nShapes * kSurEmb; k += blockDim.x) {
    const int s = k / kSurEmb, e = k % kSurEmb;
    float v = 0.f;
    for (int rr = 0; rr < kSurRows; ++rr)
      if (sSlot[rr] == s) v += sEmb[rr][e];
    G[kSurNParam + k] = v;
  }
}


// Swizzled positions of W_layer[k][n] in the forward and backward tile sets.
__device__ __forceinline__ void surSwizzle(int layer, int k, int n, int& fTile, int& fLane, int& bTile, int& bLane,
                                           int& i) {
  const int KB = layer == 0 ? kSurKB1 : 4;
  fTile = kSurTileBase[layer] + (n >> 4) * KB + (k >> 4);
  fLane = (n & 15) + 16 * ((k & 15) >> 2);
  bTile = kSurTileBase[layer] + (k >> 4) * 4 + (n >> 4);
  bLane = (k & 15) + 16 * ((n & 15) >> 2);
  i = -1;  // per operand below
}

__device__ __forceinline__ void surStoreWeight(sur_v4s* wF, sur_v4s* wB, int layer, int k, int n, float value) {
  int fT, fL, bT, bL, unused;
  surSwizzle(layer, k, n, fT, fL, bT, bL, unused);
  short* f = reinterpret_cast<short*>(wF + fT * 64 + fL);
  short* b = reinterpret_cast<short*>(wB + bT * 64 + bL);
  const short h = short(surBf16(value));
  f[k & 3] = h;
  b[n & 3] = h;
}

__global__ void k_sur_adam(SurAdamArgs A) {
  const int nAll = kSurNParam + A.nShapes * kSurEmb;
  const int stride = nAll + 1;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nAll; p += gridDim.x * blockDim.x) {
    float g = 0.f;
    for (int b = 0; b < A.blocks; ++b) g += A.partial[(long long)b * stride + p];
    const float m = A.b1 * A.m[p] + (1.f - A.b1) * g;
    const float v = A.b2 * A.v[p] + (1.f - A.b2) * g * g;
    A.m[p] = m;
    A.v[p] = v;
    float* w = p < kSurNParam ? &A.param[p] : &A.embF[p - kSurNParam];
    const float nw = *w - A.lr * (m * A.c1) / (sqrtf(v * A.c2) + A.eps);
    *w = nw;
    if (p >= kSurNParam) {
      A.emb[p - kSurNParam] = surBf16(nw);
    } else {
      for (int layer = 0; layer < 4; ++layer) {
        const int K = layer == 0 ? kSurIn : kSurHid;
        if (p >= kSurOffW[layer] && p < kSurOffW[layer] + K * kSurHid) {
          const int k = (p - kSurOffW[layer]) / kSurHid, n = (p - kSurOffW[layer]) % kSurHid;
          surStoreWeight(A.wF, A.wB, layer, k, n, nw);
        }
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float l = 0.f;
    for (int b = 0; b < A.blocks; ++b) l += A.partial[(long long)b * stride + nAll];
    *A.loss = l;
  }
}

// Fill the swizzled BF16 copies from the FP32 masters (after set_params).
__global__ void k_sur_pack(const float* param, const float* embF, int nShapes, sur_v4s* wF, sur_v4s* wB,
                           uint16_t* emb) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kSurNParam + nShapes * kSurEmb;
       p += gridDim.x * blockDim.x) {
    if (p >= kSurNParam) {
      emb[p - kSurNParam] = surBf16(embF[p - kSurNParam]);
      continue;
    }
    for (int layer = 0; layer < 4; ++layer) {
      const int K = layer == 0 ? kSurIn : kSurHid;
      if (p >= kSurOffW[layer] && p < kSurOffW[layer] + K * kSurHid)
        surStoreWeight(wF, wB, layer, (p - kSurOffW[layer]) / kSurHid, (p - kSurOffW[layer]) % kSurHid, param[p]);
    }
  }
}

// Outputs (normalised units) of a sample set and error statistics against its targets:
// stat[0] = max |out - t| (float bits), stat[1..] unused.
__global__ __launch_bounds__(256) void k_sur_eval(SurNet N, SurSamples S, float* out, unsigned int* maxErrBits, float* sumAbs) {
  const int lane = threadIdx.x & 63;
  const long long nWaves = (long long)gridDim.x * (blockDim.x >> 6);
  const long long nGroups = (S.n + 63) / 64;
  for (long long grp = blockIdx.x * (long long)(blockDim.x >> 6) + (threadIdx.x >> 6); grp < nGroups; grp += nWaves) {
    const long long i = grp * 64 + lane;
    const bool ok = i < S.n;
    const float px = ok ? S.x[3 * i] : 0.f, py = ok ? S.x[3 * i + 1] : 0.f, pz = ok ? S.x[3 * i + 2] : 0.f;
    const int slot = ok ? S.slot[i] : -1;
    const float o = surInfer64(N, px * N.invLen, py * N.invLen, pz * N.invLen, slot, lane);
    if (ok) {
      if (out) out[i] = o;
      if (maxErrBits) {
        const float e = fabsf(o - S.t[i]);
        atomicMax(maxErrBits, __float_as_uint(e));
        atomicAdd(sumAbs, e);
      }
    }
  }
}

}  // namespace lsd
