// Tutorial helpers for external-force hooks (declared in include/globalDamping.h).
#include <hip/hip_runtime.h>

#include "../../include/globalDamping.h"
#include "dmath.h"

namespace {

using lsd::d3;

__device__ inline d3 ld(const double3& v) { return lsd::mk3(v.x, v.y, v.z); }
__device__ inline void st(double3& o, d3 v) { o.x = v.x; o.y = v.y; o.z = v.z; }

__global__ void dampingKernel(double3* force, double3* torque, const double3* velocity, const double3* angularVelocity,
                              double c, size_t num) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= num) return;
  const d3 f = ld(force[i]), t = ld(torque[i]);
  st(force[i], f - (c * lsd::len(f)) * lsd::normalize(ld(velocity[i])));
  st(torque[i], t - (c * lsd::len(t)) * lsd::normalize(ld(angularVelocity[i])));
}

__global__ void constantKernel(double3* force, double3 f, double3 t, size_t index) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  d3 v = ld(force[index]);
  v = v + ld(f);
  v = v + ld(t);
  st(force[index], v);
}

}  // namespace

extern "C" void launchAddGlobalDampingForceTorque(double3* force, double3* torque, const double3* velocity,
                                                  const double3* angularVelocity, const double dampingCoefficient,
                                                  const size_t num, const size_t gridD, const size_t blockD,
                                                  hipStream_t stream) {
  if (num == 0 || gridD == 0 || blockD == 0) return;
  hipLaunchKernelGGL(dampingKernel, dim3(unsigned(gridD)), dim3(unsigned(blockD)), 0, stream, force, torque, velocity,
                     angularVelocity, dampingCoefficient, num);
}

extern "C" void launchAddGlobalConstantForceTorque(double3* force_p, double3* torque_p, const double3 constantForce,
                                                   const double3 constantTorque, const size_t index,
                                                   hipStream_t stream) {
  (void)torque_p;
  hipLaunchKernelGGL(constantKernel, dim3(1), dim3(1), 0, stream, force_p, constantForce, constantTorque, index);
}
