// HIP counterparts of the tutorial helpers FunDEM/apps/tutorial/globalDamping.cuh:1-24, which
// apps call from their external-force hook (LSSolver::addExternalForceTorque) with the device
// views of the particles.  Built into liblsdem.so.  Same names, arguments and behaviour; the
// stream argument is a hipStream_t (0 = null stream).
#pragma once
#include <cstddef>
#include <hip/hip_runtime.h>

// force -= c |force| v/|v| and torque -= c |torque| w/|w| per particle (zero vectors stay zero).
extern "C" void launchAddGlobalDampingForceTorque(double3* force, double3* torque, const double3* velocity,
                                                  const double3* angularVelocity, const double dampingCoefficient,
                                                  const size_t num, const size_t gridD, const size_t blockD,
                                                  hipStream_t stream);

// Adds constantForce and constantTorque to force[index].  The reference adds the torque to the
// force as well (globalDamping.cu:48-49); that behaviour is kept so results match.
extern "C" void launchAddGlobalConstantForceTorque(double3* force_p, double3* torque_p, const double3 constantForce,
                                                   const double3 constantTorque, const size_t index,
                                                   hipStream_t stream);
