/*
 * Affine view transform: C ABI of libNativeSPIMReconstructionCuda.so
 * (MI355X / gfx950 build).  Drop-in for fmvd_transform.h:4-24 of
 * bene51/gpu_deconvolution (C linkage here).
 *
 * The 12 floats of h_inverse are a row-major 3x4 matrix mapping target voxel
 * (x, y, z, 1) to source coordinates.  Each target voxel is the trilinear
 * interpolation of the source (zero outside), evaluated with FP64
 * coordinates, scaled by 2^B/(2^B-1) and rounded as (SAMPLE)(v*2^B + 0.5)
 * (fmvd_transform.impl.cu:41-53).
 */
#ifndef FMVD_TRANSFORM_H
#define FMVD_TRANSFORM_H

#ifdef __cplusplus
extern "C" {
#endif

/* fmvd_transform.h:4-24 / fmvd_transform.impl.cu:77-224.  h_data holds d
 * plane pointers of w*h samples.  Target planes are appended to `outfile`
 * (raw, z order).  With createTransformedMasks != 0 the float border-weight
 * plane at z = td/2 is written to `maskfile` (fmvd_transform.cu:19-67). */
void transform_cuda_8(unsigned char **h_data, int w, int h, int d, int tw, int th, int td,
                      float *h_inverse, const char *outfile, int createTransformedMasks,
                      int border, float zspacing, const char *maskfile);
void transform_cuda_16(unsigned short **h_data, int w, int h, int d, int tw, int th, int td,
                       float *h_inverse, const char *outfile, int createTransformedMasks,
                       int border, float zspacing, const char *maskfile);

/* Extensions with no counterpart in the reference: GPU-resident volumes.
 * d_src: d planes of w*h samples (contiguous) in device memory; d_dst: td
 * planes of tw*th samples.  Enqueued on `stream` (hipStream_t, NULL = null
 * stream); returns 0 on success. */
int fmvd_transform_device_8(const unsigned char *d_src, int w, int h, int d, unsigned char *d_dst,
                            int tw, int th, int td, const float *h_inverse, void *stream);
int fmvd_transform_device_16(const unsigned short *d_src, int w, int h, int d,
                             unsigned short *d_dst, int tw, int th, int td,
                             const float *h_inverse, void *stream);
/* Border-weight mask plane at z = td/2 into d_mask (tw*th floats, device). */
int fmvd_transform_mask_device(int w, int h, int d, int tw, int th, int td,
                               const float *h_inverse, int border, float zspacing, float *d_mask,
                               void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FMVD_TRANSFORM_H */
