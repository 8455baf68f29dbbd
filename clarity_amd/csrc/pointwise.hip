// Pointwise, reduction and pad/clip kernels.  Floating-point contraction is
// off in every kernel here so that each product and sum rounds once, as the
// reference's expressions do (e.g. include/internal/Complex.h:95-102); the
// results are then bit-identical to the reference's float32 arithmetic.
// Reductions accumulate in float64 over a fixed element-to-thread map and a
// fixed tree, so they are deterministic.
#include "common.hpp"

namespace clarity {

namespace {

constexpr int kBlock = 256;

unsigned blocks_for(long long n, int per_thread = 4) {
  long long b = (n + (long long)kBlock * per_thread - 1) / ((long long)kBlock * per_thread);
  if (b < 1) b = 1;
  if (b > 65535LL * 16) b = 65535LL * 16;
  return (unsigned)b;
}

#define GRID_STRIDE(i, n)                                                    \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (long long)gridDim.x * blockDim.x)

// conj_b: the product with conj(b) = (c, -d) in the same operation order
// (a correlation instead of a convolution; the negation is exact)
__global__ void modulate_kernel(const float2 *__restrict__ a,
                                const float2 *__restrict__ b, float2 *out,
                                long long n, float scale, int conj_b) {
#pragma clang fp contract(off)
  GRID_STRIDE(i, n) {
    float2 x = a[i], y = b[i];
    if (conj_b) y.y = -y.y;
    float re = x.x * y.x - x.y * y.y;
    float im = x.x * y.y + x.y * y.x;
    out[i] = make_float2(scale * re, scale * im);
  }
}

// reference src/WienerDeconvolve.cxx:48-66: Hc = conj(H) * (1/(|H|^2+eps)),
// out = scale * (Hc (x) In)
__global__ void wiener_kernel(const float2 *__restrict__ in,
                              const float2 *__restrict__ psf, float2 *out,
                              long long n, float scale, float eps) {
#pragma clang fp contract(off)
  GRID_STRIDE(i, n) {
    float2 h = psf[i];
    float mag = h.x * h.x + h.y * h.y;
    float s = 1.0f / (mag + eps);
    float cr = h.x * s;
    float ci = (-h.y) * s;
    float2 v = in[i];
    float re = cr * v.x - ci * v.y;
    float im = cr * v.y + ci * v.x;
    out[i] = make_float2(scale * re, scale * im);
  }
}

// reference src/ComputePrimitives.cxx:143-148: the float divisor's magnitude
// is compared with the double constant 1e-5, so the compare is done in double.
__global__ void divide_kernel(float *r, const float *__restrict__ a,
                              const float *__restrict__ b, float value,
                              long long n) {
  GRID_STRIDE(i, n) {
    float d = b[i];
    r[i] = ((double)fabsf(d) < 1e-5) ? value : a[i] / d;
  }
}

__global__ void guarded_divide_kernel(float *r, const float *__restrict__ a,
                                      const float *__restrict__ b, float value,
                                      long long n) {
  GRID_STRIDE(i, n) r[i] = guarded_ratio(a[i], b[i], value);
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

__global__ void multiply_sum_kernel(float *r, const float *__restrict__ a,
                                    const float *__restrict__ b, long long n,
                                    double *partials) {
#pragma clang fp contract(off)
  __shared__ double sh[kBlock];
  double acc = 0.0;
  GRID_STRIDE(i, n) {
    float v = a[i] * b[i];
    r[i] = v;
    acc += (double)v;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void sum_partials_kernel(const float *__restrict__ a, long long n,
                                    double *partials) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  GRID_STRIDE(i, n) acc += (double)a[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void idiv_partials_kernel(const float *__restrict__ g,
                                     const float *__restrict__ gh, long long n,
                                     double *partials) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  GRID_STRIDE(i, n) {
    double x = g[i], y = gh[i];
    double t = x > 0.0 ? x * log(x / y) : 0.0;
    acc += t + y - x;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void max_partials_kernel(const float *__restrict__ a, long long n,
                                    float *partials) {
  __shared__ float sh[kBlock];
  float m = a[0];
  GRID_STRIDE(i, n) {
    float v = a[i];
    if (v > m) m = v;
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && sh[threadIdx.x + s] > sh[threadIdx.x])
      sh[threadIdx.x] = sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// mode 0: out = (float)sum, and scale_out = num / out when num != nullptr
// mode 1: out = (float)(1.0 / sum)
__global__ void finish_sum_kernel(const double *__restrict__ partials,
                                  int nparts, float *out, const float *num,
                                  float *scale_out, int mode) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += partials[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    if (mode == 1) {
      out[0] = (float)(1.0 / s);
    } else {
      float f = (float)s;
      out[0] = f;
      if (num) scale_out[0] = num[0] / f;
    }
  }
}

__global__ void finish_max_kernel(const float *__restrict__ partials,
                                  int nparts, float *out) {
  __shared__ float sh[kBlock];
  float m = partials[0];
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
    if (partials[i] > m) m = partials[i];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && sh[threadIdx.x + s] > sh[threadIdx.x])
      sh[threadIdx.x] = sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void scale_dev_kernel(float *r, const float *__restrict__ a,
                                 long long n, const float *__restrict__ scale) {
  const float s = scale[0];
  GRID_STRIDE(i, n) r[i] = s * a[i];
}

__global__ void scale_kernel(float *r, const float *__restrict__ a,
                             long long n, float s) {
  GRID_STRIDE(i, n) r[i] = s * a[i];
}

// reference src/JansenVanCittertDeconvolve.cxx:106-112 (CPU form: a NaN
// update stays NaN, negatives clamp to 0)
__global__ void jvc_kernel(float *out, const float *__restrict__ in,
                           const float *ik, const float *__restrict__ ok,
                           long long n, float A, float invASq) {
#pragma clang fp contract(off)
  GRID_STRIDE(j, n) {
    float o = ok[j];
    float diff = o - A;
    float gamma = 1.0f - ((diff * diff) * invASq);
    float val = ik[j] + (gamma * (in[j] - o));
    if (val < 0.0f) val = 0.0f;
    out[j] = val;
  }
}

// reference src/ImagePadSpatialShift.cxx:42-66, index arithmetic kept as is
__global__ void pad_shift_kernel(float *dst, Dims dd, const float *__restrict__ src,
                                 Dims sd, int s0, int s1, int s2, float fill) {
  const long long n = dd.voxels();
  GRID_STRIDE(d, n) {
    int di = (int)(d % dd.x);
    long long t = d / dd.x;
    int dj = (int)(t % dd.y);
    int dk = (int)(t / dd.y);
    int sk = dk - s2;
    if (sk < 0) sk += dd.z;
    sk = sk % dd.z;
    int sj = dj - s1;
    if (sj < 0) sj += dd.y;
    sj = sj % dd.y;
    int si = di - s0;
    if (si < 0) si += dd.x;
    si = si % dd.x;
    bool inside = si >= 0 && si < sd.x && sj >= 0 && sj < sd.y && sk >= 0 &&
                  sk < sd.z;
    dst[d] = inside ? src[((long long)sk * sd.y + sj) * sd.x + si] : fill;
  }
}

__global__ void clip_kernel(float *dst, Dims dd, const float *__restrict__ src,
                            Dims sd) {
  const long long n = dd.voxels();
  GRID_STRIDE(d, n) {
    int di = (int)(d % dd.x);
    long long t = d / dd.x;
    int dj = (int)(t % dd.y);
    int dk = (int)(t / dd.y);
    dst[d] = src[((long long)dk * sd.y + dj) * sd.x + di];
  }
}

__global__ void fill_kernel(float *dst, long long n, float v) {
  GRID_STRIDE(i, n) dst[i] = v;
}

}  // namespace

ClarityResult_t k_modulate(const float2 *a, const float2 *b, float2 *out,
                           long long n, float scale, hipStream_t s, bool conj_b) {
  hipLaunchKernelGGL(modulate_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                     a, b, out, n, scale, conj_b ? 1 : 0);
  return launch_check();
}

ClarityResult_t k_wiener(const float2 *in, const float2 *psf, float2 *out,
                         long long n, float scale, float eps, hipStream_t s) {
  hipLaunchKernelGGL(wiener_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                     in, psf, out, n, scale, eps);
  return launch_check();
}

ClarityResult_t k_guarded_divide(float *r, const float *a, const float *b,
                                 float value, long long n, hipStream_t s) {
  hipLaunchKernelGGL(guarded_divide_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                     r, a, b, value, n);
  return launch_check();
}

ClarityResult_t k_divide(float *r, const float *a, const float *b, float value,
                         long long n, hipStream_t s) {
  hipLaunchKernelGGL(divide_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, r,
                     a, b, value, n);
  return launch_check();
}

ClarityResult_t k_multiply_sum(float *r, const float *a, const float *b,
                               long long n, double *partials, int nparts,
                               hipStream_t s) {
  hipLaunchKernelGGL(multiply_sum_kernel, dim3(nparts), dim3(kBlock), 0, s, r,
                     a, b, n, partials);
  return launch_check();
}

ClarityResult_t k_scale_dev(float *r, const float *a, long long n,
                            const float *scale, hipStream_t s) {
  hipLaunchKernelGGL(scale_dev_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                     r, a, n, scale);
  return launch_check();
}

ClarityResult_t k_scale(float *r, const float *a, long long n, float scale,
                        hipStream_t s) {
  hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, r,
                     a, n, scale);
  return launch_check();
}

ClarityResult_t k_jvc(float *out, const float *in, const float *ik,
                      const float *ok, long long n, float A, float invASq,
                      hipStream_t s) {
  hipLaunchKernelGGL(jvc_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, out,
                     in, ik, ok, n, A, invASq);
  return launch_check();
}

ClarityResult_t k_sum_partials(const float *a, long long n, double *partials,
                               int nparts, hipStream_t s) {
  hipLaunchKernelGGL(sum_partials_kernel, dim3(nparts), dim3(kBlock), 0, s, a,
                     n, partials);
  return launch_check();
}

ClarityResult_t k_idiv_partials(const float *g, const float *ghat, long long n,
                                double *partials, int nparts, hipStream_t s) {
  hipLaunchKernelGGL(idiv_partials_kernel, dim3(nparts), dim3(kBlock), 0, s, g,
                     ghat, n, partials);
  return launch_check();
}

ClarityResult_t k_max_partials(const float *a, long long n, float *partials,
                               int nparts, hipStream_t s) {
  hipLaunchKernelGGL(max_partials_kernel, dim3(nparts), dim3(kBlock), 0, s, a,
                     n, partials);
  return launch_check();
}

ClarityResult_t k_finish_sum(const double *partials, int nparts, float *out,
                             const float *num, float *scale_out,
                             hipStream_t s) {
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kBlock), 0, s, partials,
                     nparts, out, num, scale_out, 0);
  return launch_check();
}

ClarityResult_t k_finish_max(const float *partials, int nparts, float *out,
                             hipStream_t s) {
  hipLaunchKernelGGL(finish_max_kernel, dim3(1), dim3(kBlock), 0, s, partials,
                     nparts, out);
  return launch_check();
}

ClarityResult_t k_normalize_fp64(float *h, long long n, double *partials,
                                 int nparts, hipStream_t s) {
  // partials[nparts] holds the float scalar 1/sum (room reserved by caller)
  float *inv = reinterpret_cast<float *>(partials + nparts);
  CL_TRY(k_sum_partials(h, n, partials, nparts, s));
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kBlock), 0, s, partials,
                     nparts, inv, nullptr, nullptr, 1);
  CL_TRY(launch_check());
  return k_scale_dev(h, h, n, inv, s);
}

// reference Clarity_NormalizeArray (src/ComputePrimitives.cxx:197-209):
// sum in float, then scale by 1.0f/sum.  tmp holds 2 floats.
__global__ void recip_kernel(const float *a, float *r) { r[0] = 1.0f / a[0]; }

ClarityResult_t k_normalize_f32sum(float *h, long long n, double *partials,
                                   int nparts, float *tmp, hipStream_t s) {
  CL_TRY(k_sum_partials(h, n, partials, nparts, s));
  CL_TRY(k_finish_sum(partials, nparts, tmp, nullptr, nullptr, s));
  hipLaunchKernelGGL(recip_kernel, dim3(1), dim3(1), 0, s, tmp, tmp + 1);
  CL_TRY(launch_check());
  return k_scale_dev(h, h, n, tmp + 1, s);
}

ClarityResult_t k_pad_shift(float *dst, Dims dd, const float *src, Dims sd,
                            const int shift[3], float fill, hipStream_t s) {
  hipLaunchKernelGGL(pad_shift_kernel, dim3(blocks_for(dd.voxels())),
                     dim3(kBlock), 0, s, dst, dd, src, sd, shift[0], shift[1],
                     shift[2], fill);
  return launch_check();
}

ClarityResult_t k_clip(float *dst, Dims dd, const float *src, Dims sd,
                       hipStream_t s) {
  hipLaunchKernelGGL(clip_kernel, dim3(blocks_for(dd.voxels())), dim3(kBlock),
                     0, s, dst, dd, src, sd);
  return launch_check();
}

ClarityResult_t k_fill(float *dst, long long n, float v, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, dst,
                     n, v);
  return launch_check();
}

}  // namespace clarity
