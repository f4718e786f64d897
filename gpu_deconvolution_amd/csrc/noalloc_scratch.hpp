// Fixed device scratch for the allocation-free entry point (see
// noalloc_scratch.hip).
#pragma once
#include <stddef.h>

namespace fmvd {

// 128 MiB per device: the spectrum and estimate of 2 planes at FFT extent
// 2560 x 2560, 6 at 1536 x 1536.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;

// Device address of the current device's scratch array (resolved once per
// device with hipGetSymbolAddress; allocates nothing).
void* noalloc_scratch();

}  // namespace fmvd
