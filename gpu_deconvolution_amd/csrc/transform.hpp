// Internal (throwing) entry points of the affine transform and the border
// mask, for library code that chains them with RL on the device.  The C ABI
// of fmvd_transform.h wraps these with the error handler.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fmvd {

// Target voxels (x, y, z) for z in [0, td) of a tw x th x td volume from a
// w x h x d source (both contiguous, device); inv = 12 floats, row-major 3x4.
void transform_volume_u16(const uint16_t* src, int w, int h, int d, uint16_t* dst, int tw, int th, int td,
                          const float* inv, hipStream_t s);
void transform_volume_u8(const uint8_t* src, int w, int h, int d, uint8_t* dst, int tw, int th, int td,
                         const float* inv, hipStream_t s);
// Border-weight plane at z = td / 2 (tw x th floats, device).
void transform_mask(int w, int h, int d, int tw, int th, int td, const float* inv, int border, float zspacing,
                    float* out, hipStream_t s);

template <typename S> inline void transform_volume(const S* src, int w, int h, int d, S* dst, int tw, int th,
                                                   int td, const float* inv, hipStream_t s)
{
    if constexpr (sizeof(S) == 2) transform_volume_u16(src, w, h, d, dst, tw, th, td, inv, s);
    else transform_volume_u8(src, w, h, d, dst, tw, th, td, inv, s);
}

}  // namespace fmvd
