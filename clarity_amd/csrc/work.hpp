// Per-call work domain shared by the drivers: the observation and kernel
// uploaded once and padded on the device into imageDim + kernelDim
// (reference src/MaximumLikelihoodDeconvolve.cxx:271-289).
#pragma once

#include "common.hpp"

namespace clarity {

struct WorkGeom {
  Dims id, kd, wd;
  long long IV = 0, KV = 0, V = 0, C = 0;  // image, kernel, work voxels; bins
  float scale = 0.f;                        // 1/(nz*ny*nx) in float
};

struct Work : WorkGeom {
  void set_geometry(const WorkGeom &g) { static_cast<WorkGeom &>(*this) = g; }
  DevBuf obs;  // padded observation, low corner, zero fill
  DevBuf psf;  // padded kernel shifted by -k/2 cyclically
};

// Validates arguments and sizes; fills the geometry of w.
ClarityResult_t work_geometry(Work &w, const float *in, Clarity_Dim3 id,
                              const float *kernel, Clarity_Dim3 kd,
                              const float *out);

// Uploads in / kernel (host) and pads them on the device.
ClarityResult_t work_upload(Work &w, const float *in, const float *kernel,
                            hipStream_t s);

// Same from device pointers.
ClarityResult_t work_pad_device(Work &w, const float *in_d,
                                const float *kernel_d, hipStream_t s);

// out = C2R(modulate(R2C(in), otf)) over the work domain; spec is scratch.
// conj: modulate with conj(otf), i.e. the correlation with otf's volume.
ClarityResult_t convolve_otf(const Work &w, const float *in, const float2 *otf,
                             float *out, float2 *spec, hipStream_t s,
                             bool conj = false);

// Clips the low corner of a work-domain buffer and copies it to the host.
ClarityResult_t work_download(const Work &w, const float *src, float *host_out,
                              hipStream_t s);

}  // namespace clarity
