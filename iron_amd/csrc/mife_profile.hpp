// Optional per-kernel timing with HIP events on the launching stream.
#pragma once
#include <hip/hip_runtime.h>

namespace mife {
namespace prof {

enum Kernel : int {
  kCgSpmv = 0,
  kCgUpdate,
  kCgInit,
  kResidual64,
  kAssembleScalar,
  kAssembleElastic,
  kGather,
  kFiniteElasticity,
  kGroupedSpmv,
  kGroupedUpdate,
  kCgSpmv32,
  kCgUpdate32,
  kAssembleScalar32,
  kAssembleElastic32,
  kMinresSpmv,
  kGroupedReplace,
  kAssembleBatch,
  kFiniteElasticityResidual,
  kNumKernels
};

bool enabled();
// Records a start event; returns a ticket (or -1 when profiling is off).
int begin(hipStream_t s);
// Records the end event of a ticket and files it under `kernel`.
void end(int ticket, hipStream_t s, int kernel);
// Marks a recorded interval as not counted (e.g. a launch that returned at once).
void discard(int ticket);

}  // namespace prof
}  // namespace mife
