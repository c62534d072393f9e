// fmd_atomic.hpp -- device data layout of the atomic (AIREBO) engine.
#pragma once

#include <hip/hip_runtime.h>

#include "fmd_common.hpp"

namespace fmd {
struct AtomicEngine;
}  // namespace fmd
