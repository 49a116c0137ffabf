// Post-processing of the stacked image on the device: illumination
// compensation and the Laplacian filter that the reference's rtm applies on
// the host to copies of the image for its _illum.bin and _filtered.bin
// outputs (src/rtm/imaging.cpp:48-75; src/main.cpp:211-222).
//
// Both are one pass over the nx*nz image (HBM-bound, a few tens of MB): they
// run once per migration, after the last shot, on the image that is already
// resident on the device, so the image is not downloaded, processed on the
// host and discarded.  Results are bit-identical to the reference's host
// code: the maximum of S is order-independent, and every other value is the
// same float expression per cell, evaluated in the same order, with a
// correctly rounded division and no contraction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace rtm_hip {

// Partial maxima written by launch_image_post's first kernel.
constexpr int kPostMaxParts = 512;

// Scratch the launches need, in floats, for an image of n cells: the partial
// maxima only, whatever n (the FP64 stacks are rounded where they are read,
// and the filter of the compensated image compensates the values it reads).
inline size_t post_scratch_floats(size_t /*n*/) { return kPostMaxParts; }

// flags: bit 0 = I / (S + eps_relative*max(S) + 1e-30f), bit 1 = Laplacian
// high-pass (applied after bit 0 when both are set).  The source is either
// the FP32 image and illumination (image, illum) or the FP64 stacks
// (stack_i, stack_s), rounded once to float first, as the stack download
// does.  scratch holds post_scratch_floats(nx*nz) floats; out receives nx*nz
// floats; the sources are left unchanged.
hipError_t launch_image_post(const float* image, const float* illum, const double* stack_i,
                             const double* stack_s, int nx, int nz, float idx2, float idz2,
                             int flags, float eps_relative, float* scratch, float* out,
                             hipStream_t stream);

}  // namespace rtm_hip
