// Host-visible argument blocks of the surrogate training and evaluation kernels.
#pragma once

#include "device.h"
#include "surrogate.h"

namespace lsd {

constexpr int kSurRows = 64;  // samples per training block (4 waves x 16)

struct SurSamples {
  float* x;        // [n][3] unscaled local points
  int* slot;       // [n] family slot
  float* t;        // [n] clamped normalised target
  long long n;
};

struct SurFamily {
  const ShapeDev* shapes;
  const float* sdf;
  const int* famShape;  // [nShapes] shape id of each slot
  int nShapes;
  float invLen, clamp;
};

struct SurTrainArgs {
  SurNet net;
  SurSamples data;
  long long start, stride;  // sample of row r of the batch: (start + r * stride) mod n
  int batch;                // samples of the step (blocks x 64)
  float* partial;           // [blocks][kSurNParam + nShapes * 16 + 1]
};

struct SurAdamArgs {
  float* param;        // [kSurNParam] FP32 masters
  float* embF;         // [nShapes][16] FP32 masters
  float* m;            // [kSurNParam + nShapes * 16]
  float* v;
  const float* partial;
  int blocks, nShapes;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
  sur_v4s* wF;
  sur_v4s* wB;
  uint16_t* emb;
  float* loss;         // out: loss of the step
};

__global__ void k_sur_sample(SurFamily F, SurSamples S, unsigned long long seed);
__global__ void k_sur_train(SurTrainArgs A);
__global__ void k_sur_adam(SurAdamArgs A);
__global__ void k_sur_pack(const float* param, const float* embF, int nShapes, sur_v4s* wF, sur_v4s* wB,
                           uint16_t* emb);
__global__ __launch_bounds__(256) void k_sur_eval(SurNet N, SurSamples S, float* out, unsigned int* maxErrBits, float* sumAbs);

}  // namespace lsd
