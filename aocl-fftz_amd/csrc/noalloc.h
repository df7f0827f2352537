// Static device scratch of the allocation-free execute
// (aoclfftz_gpu_execute_noalloc in aoclfftz_gpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace fftz {

// Queue counters of the uneven batch (up to 64 size classes).
constexpr size_t kNoAllocHeaderBytes = 256;
// Largest work buffer a plan may need to run allocation-free: 64 MiB (a
// four-step transform of up to 2^22 complex128 or 2^23 complex64 elements).
constexpr size_t kNoAllocWorkBytes = size_t(64) << 20;

// Device address of this GPU's scratch array (header, then work area).  The
// first call for a device looks the address up with that device current;
// later calls return the cached address.
hipError_t noalloc_scratch(int device, void** ptr);

}  // namespace fftz
