// Host drivers of the public API (except Richardson-Lucy, in rl_plan.hip):
// convolution, Wiener, Jansen-van Cittert, blind ML (as written and
// alternating), the FFT entry points, pad/clip and the I-divergence.  All
// buffers stay on the device from upload to download; sums are float64 on
// the device and never come back to the host inside a call.
#include "common.hpp"
#include "rl_plan.hpp"
#include "work.hpp"

#include <cstdlib>
#include <cstring>

namespace clarity {

ClarityResult_t work_geometry(Work &w, const float *in, Clarity_Dim3 id,
                              const float *kernel, Clarity_Dim3 kd,
                              const float *out) {
  if (!in || !kernel || !out) return CLARITY_INVALID_ARGUMENT;
  if (!dims_ok(id) || !dims_ok(kd)) return CLARITY_INVALID_ARGUMENT;
  w.id = dims_of(id);
  w.kd = dims_of(kd);
  long long wx = (long long)id.x + kd.x, wy = (long long)id.y + kd.y,
            wz = (long long)id.z + kd.z;
  if (!fft_length_supported(wx) || !fft_length_supported(wy) || !fft_length_supported(wz))
    return CLARITY_FFT_FAILED;
  w.wd = Dims{(int)wx, (int)wy, (int)wz};
  w.IV = w.id.voxels();
  w.KV = w.kd.voxels();
  w.V = w.wd.voxels();
  w.C = (long long)w.wd.z * w.wd.y * (w.wd.x / 2 + 1);
  w.scale = inv_volume(w.wd.x, w.wd.y, w.wd.z);
  if (!ready()) return CLARITY_INVALID_OPERATION;
  return CLARITY_SUCCESS;
}

ClarityResult_t work_pad_device(Work &w, const float *in_d,
                                const float *kernel_d, hipStream_t s) {
  CL_TRY(w.obs.alloc(sizeof(float) * w.V));
  CL_TRY(w.psf.alloc(sizeof(float) * w.V));
  const int zero[3] = {0, 0, 0};
  const int kshift[3] = {-w.kd.x / 2, -w.kd.y / 2, -w.kd.z / 2};
  CL_TRY(k_pad_shift(w.obs.f(), w.wd, in_d, w.id, zero, 0.f, s));
  CL_TRY(k_pad_shift(w.psf.f(), w.wd, kernel_d, w.kd, kshift, 0.f, s));
  return CLARITY_SUCCESS;
}

ClarityResult_t work_upload(Work &w, const float *in, const float *kernel,
                            hipStream_t s) {
  DevBuf img, ker;
  CL_TRY(img.alloc(sizeof(float) * w.IV));
  CL_TRY(ker.alloc(sizeof(float) * w.KV));
  CL_TRY(host_to_device(img.p, in, sizeof(float) * w.IV, s));
  CL_TRY(host_to_device(ker.p, kernel, sizeof(float) * w.KV, s));
  CL_TRY(work_pad_device(w, img.f(), ker.f(), s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

ClarityResult_t convolve_otf(const Work &w, const float *in, const float2 *otf,
                             float *out, float2 *spec, hipStream_t s, bool conj) {
  CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, in, spec, s));
  CL_TRY(k_modulate(spec, otf, spec, w.C, w.scale, s, conj));
  return fft3d_c2r(w.wd.x, w.wd.y, w.wd.z, spec, out, s);
}

ClarityResult_t work_download(const Work &w, const float *src, float *host_out,
                              hipStream_t s) {
  DevBuf o;
  CL_TRY(o.alloc(sizeof(float) * w.IV));
  CL_TRY(k_clip(o.f(), w.id, src, w.wd, s));
  CL_TRY(device_to_host(host_out, o.p, sizeof(float) * w.IV, s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

namespace {

// reference Clarity_Convolve -> Clarity_ConvolveInternal
// (src/Convolve.cxx:41-115,146-184)
// One convolution (make_fast_convolve) or Wiener filter (make_fast_wiener)
// of `in` on the fused passes, the image-sized result to `out`; *done =
// false when the shapes are not the fused engine's.
ClarityResult_t single_fused(const Work &w, const float *in, const float *kernel, float *out,
                             bool wiener, float eps, hipStream_t s, bool *done) {
  *done = false;
  if (std::getenv("CLARITY_FORCE_GENERIC")) return CLARITY_SUCCESS;
  std::unique_ptr<RLEngine> eng = wiener ? make_fast_wiener(w, eps) : make_fast_convolve(w);
  if (!eng) return CLARITY_SUCCESS;
  *done = true;
  DevBuf img, ker;
  CL_TRY(img.alloc(sizeof(float) * w.IV));
  CL_TRY(ker.alloc(sizeof(float) * w.KV));
  CL_TRY(host_to_device(img.p, in, sizeof(float) * w.IV, s));
  CL_TRY(host_to_device(ker.p, kernel, sizeof(float) * w.KV, s));
  CL_TRY(eng->allocate());
  CL_TRY(eng->set_inputs(img.f(), ker.f(), s));
  CL_TRY(eng->reset(nullptr, s));
  CL_TRY(eng->iterate(1, s));
  CL_TRY(eng->estimate(img.f(), s));
  CL_TRY(device_to_host(out, img.p, sizeof(float) * w.IV, s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

ClarityResult_t convolve(float *in, Clarity_Dim3 id, float *kernel,
                         Clarity_Dim3 kd, float *out) {
  Work w;
  CL_TRY(work_geometry(w, in, id, kernel, kd, out));
  hipStream_t s = rt().stream;
  bool fused = false;
  CL_TRY(single_fused(w, in, kernel, out, false, 0.f, s, &fused));
  rt().last_engine = fused ? "fast" : "generic";
  if (fused) return CLARITY_SUCCESS;
  CL_TRY(work_upload(w, in, kernel, s));
  DevBuf a, b, res;
  CL_TRY(a.alloc(sizeof(float2) * w.C));
  CL_TRY(b.alloc(sizeof(float2) * w.C));
  CL_TRY(res.alloc(sizeof(float) * w.V));
  CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, w.obs.f(), a.c(), s));
  CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, w.psf.f(), b.c(), s));
  CL_TRY(k_modulate(a.c(), b.c(), a.c(), w.C, w.scale, s));
  CL_TRY(fft3d_c2r(w.wd.x, w.wd.y, w.wd.z, a.c(), res.f(), s));
  return work_download(w, res.f(), out, s);
}

// reference src/WienerDeconvolve.cxx:69-122,211-290
ClarityResult_t wiener(float *in, Clarity_Dim3 id, float *kernel,
                       Clarity_Dim3 kd, float *out, float eps) {
  Work w;
  CL_TRY(work_geometry(w, in, id, kernel, kd, out));
  hipStream_t s = rt().stream;
  bool fused = false;
  CL_TRY(single_fused(w, in, kernel, out, true, eps, s, &fused));
  rt().last_engine = fused ? "fast" : "generic";
  if (fused) return CLARITY_SUCCESS;
  CL_TRY(work_upload(w, in, kernel, s));
  DevBuf a, b, res;
  CL_TRY(a.alloc(sizeof(float2) * w.C));
  CL_TRY(b.alloc(sizeof(float2) * w.C));
  CL_TRY(res.alloc(sizeof(float) * w.V));
  CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, w.obs.f(), a.c(), s));
  CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, w.psf.f(), b.c(), s));
  CL_TRY(k_wiener(a.c(), b.c(), a.c(), w.C, w.scale, eps, s));
  CL_TRY(fft3d_c2r(w.wd.x, w.wd.y, w.wd.z, a.c(), res.f(), s));
  return work_download(w, res.f(), out, s);
}

// reference src/JansenVanCittertDeconvolve.cxx:122-192,311-389
ClarityResult_t jvc(float *in, Clarity_Dim3 id, float *kernel, Clarity_Dim3 kd,
                    float *out, int iterations) {
  if (iterations < 0) return CLARITY_INVALID_ARGUMENT;
  Work w;
  CL_TRY(work_geometry(w, in, id, kernel, kd, out));
  hipStream_t s = rt().stream;
  // max over the unpadded input (:324), on the device
  DevBuf img, ker, mparts, mval;
  CL_TRY(img.alloc(sizeof(float) * w.IV));
  CL_TRY(ker.alloc(sizeof(float) * w.KV));
  CL_TRY(host_to_device(img.p, in, sizeof(float) * w.IV, s));
  CL_TRY(host_to_device(ker.p, kernel, sizeof(float) * w.KV, s));
  CL_TRY(mparts.alloc(sizeof(float) * kParts));
  CL_TRY(mval.alloc(sizeof(float)));
  CL_TRY(k_max_partials(img.f(), w.IV, mparts.f(), kParts, s));
  CL_TRY(k_finish_max(mparts.f(), kParts, mval.f(), s));
  float mx = 0.f;
  CL_TRY(device_to_host(&mx, mval.p, sizeof(float), s));
  CL_TRY(work_pad_device(w, img.f(), ker.f(), s));
  CL_HIP(hipStreamSynchronize(s));
  img.release();
  const float A = 0.5f * mx;
  const float invASq = 1.0f / (A * A);
  // the fused passes when the reference domain is one of their sizes: one
  // convolution per iteration through the resident spectrum, the update in
  // the last x pass (with iterations <= 1 one iteration, as below)
  if (!std::getenv("CLARITY_FORCE_GENERIC")) {
    std::unique_ptr<RLEngine> eng = make_fast_jvc(w);
    if (eng) {
      rt().last_engine = "fast";
      DevBuf est;
      CL_TRY(est.alloc(sizeof(float) * w.V));
      CL_TRY(eng->allocate());
      CL_TRY(eng->set_inputs(w.obs.f(), ker.f(), s));
      CL_TRY(eng->set_jvc_constants(A, invASq, s));
      CL_TRY(eng->reset(nullptr, s));
      CL_TRY(eng->iterate(iterations > 1 ? iterations : 1, s));
      CL_TRY(eng->estimate(est.f(), s));
      return work_download(w, est.f(), out, s);
    }
  }
  ker.release();
  rt().last_engine = "generic";
  DevBuf otf, spec, ik, ok;
  CL_TRY(otf.alloc(sizeof(float2) * w.C));
  CL_TRY(spec.alloc(sizeof(float2) * w.C));
  CL_TRY(ik.alloc(sizeof(float) * w.V));
  CL_TRY(ok.alloc(sizeof(float) * w.V));
  CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, w.psf.f(), otf.c(), s));
  CL_TRY(convolve_otf(w, w.obs.f(), otf.c(), ok.f(), spec.c(), s));
  CL_TRY(k_jvc(ik.f(), w.obs.f(), w.obs.f(), ok.f(), w.V, A, invASq, s));
  for (int k = 1; k < iterations; k++) {
    CL_TRY(convolve_otf(w, ik.f(), otf.c(), ok.f(), spec.c(), s));
    CL_TRY(k_jvc(ik.f(), w.obs.f(), ik.f(), ok.f(), w.V, A, invASq, s));
  }
  return work_download(w, ik.f(), out, s);
}

// reference src/BlindMaximumLikelihoodDeconvolve.cxx:50-148,249-317 as
// written: h renormalised with a float sum every iteration, image updated
// without energy rescale, h otherwise unchanged.
ClarityResult_t blind_ml(float *in, Clarity_Dim3 id, float *kernel,
                         Clarity_Dim3 kd, float *out, int iterations) {
  if (iterations < 0) return CLARITY_INVALID_ARGUMENT;
  Work w;
  CL_TRY(work_geometry(w, in, id, kernel, kd, out));
  hipStream_t s = rt().stream;
  // h is never updated, so renormalising it every iteration leaves it (to
  // float rounding) as after the first: RL without energy rescale with the
  // normalised kernel, on the fused engine where it applies
  if (!std::getenv("CLARITY_FORCE_GENERIC")) {
    std::unique_ptr<RLEngine> eng = make_fast_rl_unscaled(w);
    if (eng) {
      rt().last_engine = "fast";
      DevBuf obs, h, o, parts;
      CL_TRY(obs.alloc(sizeof(float) * w.IV));
      CL_TRY(o.alloc(sizeof(float) * w.IV));
      CL_TRY(h.alloc(sizeof(float) * w.KV));
      CL_TRY(parts.alloc(sizeof(double) * (kParts + 2)));
      CL_TRY(host_to_device(obs.p, in, sizeof(float) * w.IV, s));
      CL_TRY(host_to_device(h.p, kernel, sizeof(float) * w.KV, s));
      CL_TRY(k_normalize_f32sum(h.f(), w.KV, parts.d(), kParts,
                                reinterpret_cast<float *>(parts.d() + kParts), s));
      CL_TRY(eng->allocate());
      CL_TRY(eng->set_inputs(obs.f(), h.f(), s));
      CL_TRY(eng->reset(nullptr, s));
      CL_TRY(eng->iterate(iterations, s));
      CL_TRY(eng->estimate(o.f(), s));
      CL_TRY(device_to_host(out, o.p, sizeof(float) * w.IV, s));
      CL_HIP(hipStreamSynchronize(s));
      return CLARITY_SUCCESS;
    }
  }
  rt().last_engine = "generic";
  CL_TRY(work_upload(w, in, kernel, s));
  DevBuf hft, spec, o, tmp, parts;
  CL_TRY(hft.alloc(sizeof(float2) * w.C));
  CL_TRY(spec.alloc(sizeof(float2) * w.C));
  CL_TRY(o.alloc(sizeof(float) * w.V));
  CL_TRY(tmp.alloc(sizeof(float) * w.V));
  CL_TRY(parts.alloc(sizeof(double) * (kParts + 2)));
  CL_HIP(hipMemcpyAsync(o.p, w.obs.p, sizeof(float) * w.V, hipMemcpyDeviceToDevice, s));
  float *scal = reinterpret_cast<float *>(parts.d() + kParts);
  for (int k = 0; k < iterations; k++) {
    CL_TRY(k_normalize_f32sum(w.psf.f(), w.V, parts.d(), kParts, scal, s));
    CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, w.psf.f(), hft.c(), s));
    CL_TRY(convolve_otf(w, o.f(), hft.c(), tmp.f(), spec.c(), s));
    CL_TRY(k_divide(tmp.f(), w.obs.f(), tmp.f(), 0.f, w.V, s));
    CL_TRY(convolve_otf(w, tmp.f(), hft.c(), tmp.f(), spec.c(), s));
    CL_TRY(k_multiply_sum(o.f(), o.f(), tmp.f(), w.V, parts.d(), kParts, s));
  }
  return work_download(w, o.f(), out, s);
}

// Alternating blind RL on the fused RL engines: the image steps are RL
// iterations of the kernel-fixed problem (the engine on its alias-free work
// domain), the kernel steps RL iterations of the image-fixed problem with
// the kernel as the estimate (make_fast_rl_kernel_step: reference work
// domain, the correction a correlation with the image, renormalisation to
// sum 1).  The image and kernel estimates stay in their caller layouts,
// each engine's estimate the other's blurring input.  initial (host,
// image dims) or NULL = the observation.  *done = false when the shapes are
// not the fused engines' (different or odd dimensions, unsupported sizes).
ClarityResult_t blind_rl_fused(const Work &w, const float *in, const float *kernel,
                               const float *initial, float *out, float *kernel_out,
                               int outer, int inner, hipStream_t s, bool *done) {
  *done = false;
  if (std::getenv("CLARITY_FORCE_GENERIC")) return CLARITY_SUCCESS;
  std::unique_ptr<RLEngine> img = make_fast_rl(w);
  std::unique_ptr<RLEngine> ker = make_fast_rl_kernel_step(w);
  if (!img || !ker) return CLARITY_SUCCESS;
  *done = true;
  CL_TRY(img->allocate());
  CL_TRY(ker->allocate());
  DevBuf obs, o, h, parts;
  CL_TRY(obs.alloc(sizeof(float) * w.IV));
  CL_TRY(o.alloc(sizeof(float) * w.IV));
  CL_TRY(h.alloc(sizeof(float) * w.KV));
  CL_TRY(parts.alloc(sizeof(double) * (kParts + 2)));
  CL_TRY(host_to_device(obs.p, in, sizeof(float) * w.IV, s));
  CL_TRY(host_to_device(h.p, kernel, sizeof(float) * w.KV, s));
  if (initial)
    CL_TRY(host_to_device(o.p, initial, sizeof(float) * w.IV, s));
  else
    CL_HIP(hipMemcpyAsync(o.p, obs.p, sizeof(float) * w.IV, hipMemcpyDeviceToDevice, s));
  CL_TRY(k_normalize_fp64(h.f(), w.KV, parts.d(), kParts, s));
  for (int it = 0; it < outer; it++) {
    CL_TRY(img->set_inputs(obs.f(), h.f(), s));
    CL_TRY(img->reset(o.f(), s));
    CL_TRY(img->iterate(inner, s));
    CL_TRY(img->estimate(o.f(), s));
    CL_TRY(ker->set_inputs(obs.f(), o.f(), s));
    CL_TRY(ker->reset(h.f(), s));
    CL_TRY(ker->iterate(inner, s));
    CL_TRY(ker->estimate(h.f(), s));
  }
  if (kernel_out)
    CL_TRY(device_to_host(kernel_out, h.p, sizeof(float) * w.KV, s));
  CL_TRY(device_to_host(out, o.p, sizeof(float) * w.IV, s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

// Alternating blind RL (see ClarityEx.h); initial (host, image dims) or
// NULL = the observation.
ClarityResult_t blind_rl(float *in, Clarity_Dim3 id, float *kernel,
                         Clarity_Dim3 kd, const float *initial, float *out,
                         float *kernel_out, int outer, int inner) {
  if (outer < 0 || inner < 0) return CLARITY_INVALID_ARGUMENT;
  Work w;
  CL_TRY(work_geometry(w, in, id, kernel, kd, out));
  hipStream_t s = rt().stream;
  bool fused = false;
  CL_TRY(blind_rl_fused(w, in, kernel, initial, out, kernel_out, outer, inner, s,
                        &fused));
  rt().last_engine = fused ? "fast" : "generic";
  if (fused) return CLARITY_SUCCESS;
  CL_TRY(work_upload(w, in, kernel, s));
  DevBuf hft, oft, spec, o, s1, s2, parts, scal;
  CL_TRY(hft.alloc(sizeof(float2) * w.C));
  CL_TRY(oft.alloc(sizeof(float2) * w.C));
  CL_TRY(spec.alloc(sizeof(float2) * w.C));
  CL_TRY(o.alloc(sizeof(float) * w.V));
  CL_TRY(s1.alloc(sizeof(float) * w.V));
  CL_TRY(s2.alloc(sizeof(float) * w.V));
  CL_TRY(parts.alloc(sizeof(double) * (kParts + 2)));
  CL_TRY(scal.alloc(sizeof(float) * 4));  // E, sum, scale
  float *E = scal.f(), *S = scal.f() + 1, *c = scal.f() + 2;
  float *h = w.psf.f();
  if (initial) {
    DevBuf init;
    CL_TRY(init.alloc(sizeof(float) * w.IV));
    CL_TRY(host_to_device(init.p, initial, sizeof(float) * w.IV, s));
    const int zero[3] = {0, 0, 0};
    CL_TRY(k_pad_shift(o.f(), w.wd, init.f(), w.id, zero, 0.f, s));
    CL_HIP(hipStreamSynchronize(s));
  } else {
    CL_HIP(hipMemcpyAsync(o.p, w.obs.p, sizeof(float) * w.V, hipMemcpyDeviceToDevice, s));
  }
  CL_TRY(k_normalize_fp64(h, w.V, parts.d(), kParts, s));
  CL_TRY(k_sum_partials(w.obs.f(), w.V, parts.d(), kParts, s));
  CL_TRY(k_finish_sum(parts.d(), kParts, E, nullptr, nullptr, s));
  for (int it = 0; it < outer; it++) {
    CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, h, hft.c(), s));
    for (int k = 0; k < inner; k++) {
      CL_TRY(convolve_otf(w, o.f(), hft.c(), s1.f(), spec.c(), s));
      CL_TRY(k_divide(s1.f(), w.obs.f(), s1.f(), 0.f, w.V, s));
      CL_TRY(convolve_otf(w, s1.f(), hft.c(), s2.f(), spec.c(), s));
      CL_TRY(k_multiply_sum(o.f(), o.f(), s2.f(), w.V, parts.d(), kParts, s));
      CL_TRY(k_finish_sum(parts.d(), kParts, S, E, c, s));
      CL_TRY(k_scale_dev(o.f(), o.f(), w.V, c, s));
    }
    CL_TRY(fft3d_r2c(w.wd.x, w.wd.y, w.wd.z, o.f(), oft.c(), s));
    for (int k = 0; k < inner; k++) {
      CL_TRY(convolve_otf(w, h, oft.c(), s1.f(), spec.c(), s));
      CL_TRY(k_divide(s1.f(), w.obs.f(), s1.f(), 0.f, w.V, s));
      // the correction correlates the ratio with the image estimate
      CL_TRY(convolve_otf(w, s1.f(), oft.c(), s2.f(), spec.c(), s, true));
      CL_TRY(k_multiply_sum(h, h, s2.f(), w.V, parts.d(), kParts, s));
      CL_TRY(k_normalize_fp64(h, w.V, parts.d(), kParts, s));
    }
  }
  if (kernel_out) {
    // undo the -k/2 cyclic shift, then take the kernel-sized low corner
    DevBuf un, ko;
    CL_TRY(un.alloc(sizeof(float) * w.V));
    CL_TRY(ko.alloc(sizeof(float) * w.KV));
    const int back[3] = {w.kd.x / 2, w.kd.y / 2, w.kd.z / 2};
    CL_TRY(k_pad_shift(un.f(), w.wd, h, w.wd, back, 0.f, s));
    CL_TRY(k_clip(ko.f(), w.kd, un.f(), w.wd, s));
    CL_TRY(device_to_host(kernel_out, ko.p, sizeof(float) * w.KV, s));
  }
  return work_download(w, o.f(), out, s);
}

ClarityResult_t fft_r2c_api(int nx, int ny, int nz, float *in, float *out) {
  if (!in || !out || nx <= 0 || ny <= 0 || nz <= 0) return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  hipStream_t s = rt().stream;
  const long long V = (long long)nx * ny * nz;
  const long long C = (long long)nz * ny * (nx / 2 + 1);
  DevBuf a, b;
  CL_TRY(a.alloc(sizeof(float) * V));
  CL_TRY(b.alloc(sizeof(float2) * C));
  CL_TRY(host_to_device(a.p, in, sizeof(float) * V, s));
  CL_TRY(fft3d_r2c(nx, ny, nz, a.f(), b.c(), s));
  CL_TRY(device_to_host(out, b.p, sizeof(float2) * C, s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

ClarityResult_t fft_c2r_api(int nx, int ny, int nz, float *in, float *out) {
  if (!in || !out || nx <= 0 || ny <= 0 || nz <= 0) return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  hipStream_t s = rt().stream;
  const long long V = (long long)nx * ny * nz;
  const long long C = (long long)nz * ny * (nx / 2 + 1);
  DevBuf a, b;
  CL_TRY(a.alloc(sizeof(float2) * C));
  CL_TRY(b.alloc(sizeof(float) * V));
  CL_TRY(host_to_device(a.p, in, sizeof(float2) * C, s));
  CL_TRY(fft3d_c2r(nx, ny, nz, a.c(), b.f(), s));
  CL_TRY(device_to_host(out, b.p, sizeof(float) * V, s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

ClarityResult_t pad_api(float *dst, Clarity_Dim3 dd, float *src,
                        Clarity_Dim3 sd, int shift[3], float fill) {
  // reference src/ImagePadSpatialShift.cxx:38-40: NULL -> INVALID_ARGUMENT
  if (!dst || !src || !shift) return CLARITY_INVALID_ARGUMENT;
  if (!dims_ok(dd) || !dims_ok(sd)) return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  hipStream_t s = rt().stream;
  Dims D = dims_of(dd), S = dims_of(sd);
  DevBuf a, b;
  CL_TRY(a.alloc(sizeof(float) * S.voxels()));
  CL_TRY(b.alloc(sizeof(float) * D.voxels()));
  CL_TRY(host_to_device(a.p, src, sizeof(float) * S.voxels(), s));
  CL_TRY(k_pad_shift(b.f(), D, a.f(), S, shift, fill, s));
  CL_TRY(device_to_host(dst, b.p, sizeof(float) * D.voxels(), s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

ClarityResult_t clip_api(float *dst, Clarity_Dim3 dd, float *src,
                         Clarity_Dim3 sd) {
  if (!dst || !src) return CLARITY_INVALID_ARGUMENT;  // src/ImageClip.cxx:34-36
  if (!dims_ok(dd) || !dims_ok(sd)) return CLARITY_INVALID_ARGUMENT;
  if (dd.x > sd.x || dd.y > sd.y || dd.z > sd.z) return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  hipStream_t s = rt().stream;
  Dims D = dims_of(dd), S = dims_of(sd);
  DevBuf a, b;
  CL_TRY(a.alloc(sizeof(float) * S.voxels()));
  CL_TRY(b.alloc(sizeof(float) * D.voxels()));
  CL_TRY(host_to_device(a.p, src, sizeof(float) * S.voxels(), s));
  CL_TRY(k_clip(b.f(), D, a.f(), S, s));
  CL_TRY(device_to_host(dst, b.p, sizeof(float) * D.voxels(), s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

ClarityResult_t idiv_api(const float *g, const float *ghat, int64_t n,
                         double *result) {
  if (!g || !ghat || !result || n <= 0) return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  hipStream_t s = rt().stream;
  DevBuf a, b, parts;
  CL_TRY(a.alloc(sizeof(float) * n));
  CL_TRY(b.alloc(sizeof(float) * n));
  CL_TRY(parts.alloc(sizeof(double) * kParts));
  CL_TRY(host_to_device(a.p, g, sizeof(float) * n, s));
  CL_TRY(host_to_device(b.p, ghat, sizeof(float) * n, s));
  CL_TRY(k_idiv_partials(a.f(), b.f(), n, parts.d(), kParts, s));
  std::vector<double> h(kParts);
  CL_TRY(device_to_host(h.data(), parts.p, sizeof(double) * kParts, s));
  CL_HIP(hipStreamSynchronize(s));
  double acc = 0.0;
  for (double v : h) acc += v;
  *result = acc;
  return CLARITY_SUCCESS;
}

// reference Clarity_DivideArraysComponentWise (src/ComputePrimitives.cxx:
// 121-158) computed by the fused passes' guarded ratio, on host arrays
ClarityResult_t divide_api(float *result, const float *a, const float *b,
                           float value, int64_t n) {
  if (!result || !a || !b || n <= 0) return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  hipStream_t s = rt().stream;
  DevBuf da, db, dr;
  CL_TRY(da.alloc(sizeof(float) * n));
  CL_TRY(db.alloc(sizeof(float) * n));
  CL_TRY(dr.alloc(sizeof(float) * n));
  CL_TRY(host_to_device(da.p, a, sizeof(float) * n, s));
  CL_TRY(host_to_device(db.p, b, sizeof(float) * n, s));
  CL_TRY(k_guarded_divide(dr.f(), da.f(), db.f(), value, n, s));
  CL_TRY(device_to_host(result, dr.p, sizeof(float) * n, s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

// Clarity_Convolve on device buffers without allocating: the generic passes
// over the reference's work domain (image + kernel per axis, kernel
// centred by the -k/2 cyclic shift, src/Convolve.cxx:41-115,146-184), every
// intermediate (the padded volume, both spectra, the three twiddle tables)
// carved out of the per-device scratch array, all on one stream with no
// host synchronisation and no memset (every region is written in full
// before it is read).  Calls that overlap on one GPU share the array and
// race.
ClarityResult_t convolve_noalloc(const float *in, Clarity_Dim3 id, const float *kernel,
                                 Clarity_Dim3 kd, float *out, void *stream) {
  if (!in || !kernel || !out) return CLARITY_INVALID_ARGUMENT;
  if (!dims_ok(id) || !dims_ok(kd)) return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  Runtime &r = rt();
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : r.stream;
  auto it = r.noalloc_scratch.find(r.device);
  if (it == r.noalloc_scratch.end() || !it->second) return CLARITY_INVALID_OPERATION;
  unsigned char *base = static_cast<unsigned char *>(it->second);
  const Dims I = dims_of(id), K = dims_of(kd);
  const long long wx = (long long)I.x + K.x, wy = (long long)I.y + K.y,
                  wz = (long long)I.z + K.z;
  if (wx > kMaxGenericN || wy > kMaxGenericN || wz > kMaxGenericN)
    return CLARITY_INVALID_ARGUMENT;
  const Dims W{(int)wx, (int)wy, (int)wz};
  AxisPlan px, py, pz;
  CL_TRY(axis_factor(W.x, &px));
  CL_TRY(axis_factor(W.y, &py));
  CL_TRY(axis_factor(W.z, &pz));
  // scratch layout, 256-byte aligned regions: padded real volume P, spectra
  // A and B, twiddle tables for x, y, z
  const long long V = W.voxels();
  const long long C = (long long)W.z * W.y * (W.x / 2 + 1);
  auto up = [](unsigned long long b) { return (b + 255ull) & ~255ull; };
  const unsigned long long oP = 0, oA = oP + up(4ull * V), oB = oA + up(8ull * C),
                           oX = oB + up(8ull * C), oY = oX + up(8ull * W.x),
                           oZ = oY + up(8ull * W.y), end = oZ + up(8ull * W.z);
  if (end > kNoAllocScratchBytes) return CLARITY_INVALID_ARGUMENT;
  float *P = reinterpret_cast<float *>(base + oP);
  float2 *A = reinterpret_cast<float2 *>(base + oA);
  float2 *B = reinterpret_cast<float2 *>(base + oB);
  px.tw = reinterpret_cast<float2 *>(base + oX);
  py.tw = reinterpret_cast<float2 *>(base + oY);
  pz.tw = reinterpret_cast<float2 *>(base + oZ);
  CL_TRY(k_twiddles(reinterpret_cast<float2 *>(base + oX), W.x, s));
  CL_TRY(k_twiddles(reinterpret_cast<float2 *>(base + oY), W.y, s));
  CL_TRY(k_twiddles(reinterpret_cast<float2 *>(base + oZ), W.z, s));
  const int zero[3] = {0, 0, 0};
  const int kshift[3] = {-K.x / 2, -K.y / 2, -K.z / 2};
  CL_TRY(k_pad_shift(P, W, in, I, zero, 0.f, s));
  CL_TRY(fft3d_r2c_planned(px, py, pz, P, A, s));
  CL_TRY(k_pad_shift(P, W, kernel, K, kshift, 0.f, s));
  CL_TRY(fft3d_r2c_planned(px, py, pz, P, B, s));
  CL_TRY(k_modulate(A, B, A, C, inv_volume(W.x, W.y, W.z), s));
  CL_TRY(fft3d_c2r_planned(px, py, pz, A, P, s));
  CL_TRY(k_clip(out, I, P, W, s));
  r.last_engine = "generic";
  return CLARITY_SUCCESS;
}

// Zero-filling stubs kept for ABI completeness (reference
// src/SmoothedJansenVanCittertDeconvolve.cxx:28-39,
// src/IDivergenceDeconvolve.cxx:59-72).
ClarityResult_t zero_stub(float *out, Clarity_Dim3 id) {
  if (!out || !dims_ok(id)) return CLARITY_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(float) * (size_t)dims_of(id).voxels());
  return CLARITY_SUCCESS;
}

}  // namespace
}  // namespace clarity

using namespace clarity;

#define API_LOCK std::lock_guard<std::recursive_mutex> lk_(rt().mu)

extern "C" {

ClarityResult_t Clarity_Convolve(float *inImage, Clarity_Dim3 imageDim,
                                 float *kernel, Clarity_Dim3 kernelDim,
                                 float *outImage) {
  API_LOCK;
  return convolve(inImage, imageDim, kernel, kernelDim, outImage);
}

ClarityResult_t Clarity_WienerDeconvolve(float *inImage, Clarity_Dim3 imageDim,
                                         float *kernelImage,
                                         Clarity_Dim3 kernelDim,
                                         float *outImage, float epsilon) {
  API_LOCK;
  return wiener(inImage, imageDim, kernelImage, kernelDim, outImage, epsilon);
}

ClarityResult_t Clarity_JansenVanCittertDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage, int iterations) {
  API_LOCK;
  return jvc(inImage, imageDim, kernelImage, kernelDim, outImage, iterations);
}

ClarityResult_t Clarity_BlindMaximumLikelihoodDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage, int iterations) {
  API_LOCK;
  return blind_ml(inImage, imageDim, kernelImage, kernelDim, outImage,
                  iterations);
}

ClarityResult_t ClarityEx_BlindRL(float *inImage, Clarity_Dim3 imageDim,
                                  float *kernelImage, Clarity_Dim3 kernelDim,
                                  float *outImage, float *kernelOut, int outer,
                                  int inner) {
  API_LOCK;
  return blind_rl(inImage, imageDim, kernelImage, kernelDim, nullptr, outImage,
                  kernelOut, outer, inner);
}

ClarityResult_t ClarityEx_BlindRLFrom(float *inImage, Clarity_Dim3 imageDim,
                                      float *kernelImage, Clarity_Dim3 kernelDim,
                                      float *initialImage, float *outImage,
                                      float *kernelOut, int outer, int inner) {
  API_LOCK;
  return blind_rl(inImage, imageDim, kernelImage, kernelDim, initialImage,
                  outImage, kernelOut, outer, inner);
}

ClarityResult_t Clarity_SmoothedJansenVanCittertDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage, unsigned iterations,
    unsigned smoothInterval, float *smoothSigma[3]) {
  (void)inImage; (void)kernelImage; (void)kernelDim; (void)iterations;
  (void)smoothInterval; (void)smoothSigma;
  return zero_stub(outImage, imageDim);
}

ClarityResult_t Clarity_IDivergenceDeconvolve(float *inImage,
                                              Clarity_Dim3 imageDim,
                                              float *kernelImage,
                                              Clarity_Dim3 kernelDim,
                                              float *outImage) {
  (void)inImage; (void)kernelImage; (void)kernelDim;
  return zero_stub(outImage, imageDim);
}

ClarityResult_t Clarity_FFT_R2C_float(int nx, int ny, int nz, float *in,
                                      float *out) {
  API_LOCK;
  return fft_r2c_api(nx, ny, nz, in, out);
}

ClarityResult_t Clarity_FFT_C2R_float(int nx, int ny, int nz, float *in,
                                      float *out) {
  API_LOCK;
  return fft_c2r_api(nx, ny, nz, in, out);
}

ClarityResult_t Clarity_ImagePadSpatialShift(float *dst, Clarity_Dim3 dstDim,
                                             float *src, Clarity_Dim3 srcDim,
                                             int shift[3], float fillValue) {
  API_LOCK;
  return pad_api(dst, dstDim, src, srcDim, shift, fillValue);
}

ClarityResult_t Clarity_ImageClip(float *dst, Clarity_Dim3 dstDim, float *src,
                                  Clarity_Dim3 srcDim) {
  API_LOCK;
  return clip_api(dst, dstDim, src, srcDim);
}

ClarityResult_t ClarityEx_IDivergence(const float *g, const float *ghat,
                                      int64_t n, double *result) {
  API_LOCK;
  return idiv_api(g, ghat, n, result);
}

ClarityResult_t ClarityEx_ConvolveNoAlloc(const float *inImage, Clarity_Dim3 imageDim,
                                          const float *kernelImage,
                                          Clarity_Dim3 kernelDim, float *outImage,
                                          void *stream) {
  API_LOCK;
  return convolve_noalloc(inImage, imageDim, kernelImage, kernelDim, outImage, stream);
}

ClarityResult_t ClarityEx_DivideArraysComponentWise(float *result, const float *a,
                                                    const float *b, float value,
                                                    int64_t n) {
  API_LOCK;
  return divide_api(result, a, b, value, n);
}

}  // extern "C"
