// Shared declarations of the HIP Clarity library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/Clarity.h"
#include "../../include/ClarityEx.h"

namespace clarity {

// ---------------------------------------------------------------------------
// Runtime state (one process, one device)
// ---------------------------------------------------------------------------

struct Runtime {
  std::recursive_mutex mu;
  int refcount = 0;
  int device = -1;
  hipStream_t stream = nullptr;
  std::string device_name;
  unsigned host_threads = 1;
  long long alloc_count = 0;       // device allocations since Register
  long long fault_every = 0;       // fail every k-th allocation (tests)
  std::multimap<size_t, void *> cache;  // freed device blocks by size
  size_t cache_bytes = 0;
  std::map<int, float2 *> twiddles;     // w_N^j tables, j < N
  std::map<long long, float2 *> tables; // four-step and Bluestein tables
  const char *last_engine = "";         // engine of the last driver call
  // device address of the allocation-free entry point's scratch array, per
  // device (looked up at registration; nullptr if the lookup failed)
  std::map<int, void *> noalloc_scratch;
  // pinned staging buffers of host_to_device / device_to_host, each with
  // the event of its last DMA (allocated on first use)
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  int stage_next = 0;
};

Runtime &rt();

// A registration reference held by a long-lived handle (RL plan, network):
// the runtime's stream, twiddle tables and memory cache, which the handle's
// kernels use, stay alive until the handle is destroyed, even after the
// caller's last Clarity_UnRegister.  Declare it as the handle's first
// member, so that it is released after the handle's device buffers.
struct RegHold {
  bool held = false;
  RegHold() = default;
  RegHold(const RegHold &) = delete;
  RegHold &operator=(const RegHold &) = delete;
  ClarityResult_t acquire() {
    if (held) return CLARITY_SUCCESS;
    const ClarityResult_t r = Clarity_Register();
    held = r == CLARITY_SUCCESS;
    return r;
  }
  ~RegHold() {
    if (held) (void)Clarity_UnRegister();
  }
};

// True once Clarity_Register succeeded (and before the last UnRegister).
bool ready();

// Maps a HIP error to the library's result code.
ClarityResult_t from_hip(hipError_t e);

#define CL_HIP(expr)                                        \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return ::clarity::from_hip(e_);   \
  } while (0)

#define CL_TRY(expr)                                        \
  do {                                                      \
    ClarityResult_t r_ = (expr);                            \
    if (r_ != CLARITY_SUCCESS) return r_;                   \
  } while (0)

// Checks the last kernel launch.
ClarityResult_t launch_check();

// ---------------------------------------------------------------------------
// Device memory: a size-keyed cache of freed blocks in front of hipMalloc so
// that repeated calls of one shape do not allocate.  Blocks return to the
// cache on release and are freed at the last UnRegister.
// ---------------------------------------------------------------------------

// Copies between caller (pageable) host memory and the device through two
// pinned staging buffers, the host side split over threads, the DMA of one
// chunk overlapping the host copy of the next (about 50 GB/s each way on
// MI355X against 17-28 GB/s for hipMemcpy from pageable memory).
// host_to_device returns once src has been read (the last DMA may still run
// on s, ahead of later work on s); device_to_host returns with dst filled,
// after the work queued on s before it.  Copies under 4 MB go directly.
ClarityResult_t host_to_device(void *dst, const void *src, size_t bytes, hipStream_t s);
ClarityResult_t device_to_host(void *dst, const void *src, size_t bytes, hipStream_t s);

ClarityResult_t dev_alloc(void **p, size_t bytes);
void dev_release(void *p, size_t bytes);
void dev_cache_clear();

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  ClarityResult_t alloc(size_t n) {
    release();
    ClarityResult_t r = dev_alloc(&p, n);
    if (r == CLARITY_SUCCESS) bytes = n;
    return r;
  }
  void release() {
    if (p) dev_release(p, bytes);
    p = nullptr;
    bytes = 0;
  }
  float *f() const { return static_cast<float *>(p); }
  float2 *c() const { return static_cast<float2 *>(p); }
  double *d() const { return static_cast<double *>(p); }
};

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------

struct Dims {
  int x = 0, y = 0, z = 0;
  __host__ __device__ long long voxels() const { return (long long)x * y * z; }
};

inline Dims dims_of(Clarity_Dim3 d) { return Dims{d.x, d.y, d.z}; }
inline bool dims_ok(Clarity_Dim3 d) { return d.x > 0 && d.y > 0 && d.z > 0; }

// The reference's divide guard compares a float |d| with the double 1e-5
// (src/ComputePrimitives.cxx:143-148).  No float lies in [1e-5, 0x1.4f8b5ap-17)
// (0x1.4f8b58p-17 < 1e-5 < 0x1.4f8b5ap-17 are neighbours), so the same test
// in float is |d| < kDivGuard.
constexpr float kDivGuard = 0x1.4f8b5ap-17f;

// a / b for normal b (the guarded ratio has |b| >= 1e-5): the hardware
// reciprocal (1 ulp) and one Newton correction of the quotient, q + (a - b
// q) / b with the residual exact in one fma; within 1 ulp of the correctly
// rounded quotient for 1e-5 <= |b| < 2^126 in 4 instructions, where IEEE
// division takes about 10 (scaling, reciprocal, refinement, fix-up for b
// near the range limits).
__device__ __forceinline__ float quotient(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  return fmaf(fmaf(-b, q, a), r, q);
}

// The divide guard of the fused passes: `value` where |b| < 1e-5 (the
// reference's double compare, src/ComputePrimitives.cxx:143-148), else
// quotient(a, b).
__device__ __forceinline__ float guarded_ratio(float a, float b, float value) {
  return (fabsf(b) < kDivGuard) ? value : quotient(a, b);
}

// 1/(nz*ny*nx) as the reference computes it: ((float)nz * ny) * nx in float
// (reference src/Convolve.cxx:199).
inline float inv_volume(int nx, int ny, int nz) {
  float v = (float)nz;
  v = v * (float)ny;
  v = v * (float)nx;
  return 1.0f / v;
}

// ---------------------------------------------------------------------------
// Generic FFT (fft_generic.hip)
// ---------------------------------------------------------------------------

constexpr int kMaxStages = 24;
constexpr int kMaxGenericN = 4096;  // LDS: two ping-pong tiles of N*TP float2

struct AxisPlan {
  int N = 0;
  int nst = 0;
  int radix[kMaxStages] = {};
  const float2 *tw = nullptr;  // device table w_N^j = exp(-2 pi i j/N)
};

// Pencil (a, b) has element j at a*sa + b*sb + j*se (complex or real units).
// b is the index tiled TP at a time, Lin the number of leading nonzero input
// elements, Lout the number of leading outputs written.
struct PencilGeom {
  long long nA = 1, nB = 1;
  long long ia = 0, ib = 0, ie = 1;
  long long oa = 0, ob = 0, oe = 1;
  int Lin = 0, Lout = 0;
};

ClarityResult_t axis_plan(int n, AxisPlan *out);
// The factorisation of axis_plan without the twiddle table (tw = nullptr).
ClarityResult_t axis_factor(int n, AxisPlan *out);
// tw[j] = w_n^j for j < n, computed on the device in float64 and rounded to
// float as axis_plan's host table is.
ClarityResult_t k_twiddles(float2 *tw, int n, hipStream_t s);

// Complex pencils, in may equal out.  sign -1 forward, +1 inverse.
ClarityResult_t fft_c2c_pass(const float2 *in, float2 *out, const AxisPlan &ap,
                             const PencilGeom &g, int sign, hipStream_t s);
// Real rows (length Lin, zero padded to N) -> first N/2+1 bins.
ClarityResult_t fft_r2c_rows(const float *in, float2 *out, const AxisPlan &ap,
                             const PencilGeom &g, hipStream_t s);
// Half-complex rows (N/2+1 bins) -> first Lout real outputs.
ClarityResult_t fft_c2r_rows(const float2 *in, float *out, const AxisPlan &ap,
                             const PencilGeom &g, hipStream_t s);

// Whole-volume transforms in FFTW layout (z, y, kx), device pointers.
ClarityResult_t fft3d_r2c(int nx, int ny, int nz, const float *in, float2 *out,
                          hipStream_t s);
// spec is used as scratch (overwritten), like FFTW's c2r.
ClarityResult_t fft3d_c2r(int nx, int ny, int nz, float2 *spec, float *out,
                          hipStream_t s);
// Axis lengths the 3D transforms accept: every n <= kMaxFftLength.  Lengths
// the LDS passes take in one tile (<= kMaxGenericN, prime factors <= 64)
// run there; the others by four-step or Bluestein (fft_large.hip).
constexpr long long kMaxFftLength = 1LL << 23;
bool fft_length_supported(long long n);
bool fft_length_direct(int n);
// What fft_axis_large gathers: complex elements, real elements (imaginary
// part 0), or a half spectrum of n/2 + 1 bins extended by Hermitian
// symmetry (the imaginary parts of bin 0 and, for even n, bin n/2 ignored,
// as a c2r does).
enum FftGather { kGatherComplex = 0, kGatherReal = 1, kGatherHalf = 2 };
// One axis of length n of any length over the pencils of g (input element
// m of pencil (a, b) at a ia + b ib + m ie, output element k at a oa + b ob
// + k oe): gather, transform with `sign`, scatter the first `count` outputs
// (their real parts when real_out).  in may equal out.  Synchronises s.
ClarityResult_t fft_axis_large(const void *in, void *out, const PencilGeom &g, int n,
                               int kind, int sign, int count, bool real_out, hipStream_t s);

// The same with the axes' plans (twiddle tables included) given.
ClarityResult_t fft3d_r2c_planned(const AxisPlan &px, const AxisPlan &py,
                                  const AxisPlan &pz, const float *in, float2 *out,
                                  hipStream_t s);
ClarityResult_t fft3d_c2r_planned(const AxisPlan &px, const AxisPlan &py,
                                  const AxisPlan &pz, float2 *spec, float *out,
                                  hipStream_t s);

// Size of the per-device scratch array of the allocation-free entry point
// (ClarityEx_ConvolveNoAlloc), a static device array in runtime.hip.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;

// ---------------------------------------------------------------------------
// Pointwise kernels (pointwise.hip), all on stream s, n elements.
// ---------------------------------------------------------------------------

// out = scale * a * b (conj_b: scale * a * conj(b), a correlation)
ClarityResult_t k_modulate(const float2 *a, const float2 *b, float2 *out,
                           long long n, float scale, hipStream_t s,
                           bool conj_b = false);
ClarityResult_t k_wiener(const float2 *in, const float2 *psf, float2 *out,
                         long long n, float scale, float eps, hipStream_t s);
// r = guarded_ratio(a, b, value): the fused passes' divide
ClarityResult_t k_guarded_divide(float *r, const float *a, const float *b,
                                 float value, long long n, hipStream_t s);
ClarityResult_t k_divide(float *r, const float *a, const float *b,
                         float value, long long n, hipStream_t s);
// r = a * b, with float64 per-block partial sums of r into partials.
ClarityResult_t k_multiply_sum(float *r, const float *a, const float *b,
                               long long n, double *partials, int nparts,
                               hipStream_t s);
ClarityResult_t k_scale_dev(float *r, const float *a, long long n,
                            const float *scale, hipStream_t s);
ClarityResult_t k_scale(float *r, const float *a, long long n, float scale,
                        hipStream_t s);
ClarityResult_t k_jvc(float *out, const float *in, const float *ik,
                      const float *ok, long long n, float A, float invASq,
                      hipStream_t s);
// float64 partial sums of a into partials[nparts]
ClarityResult_t k_sum_partials(const float *a, long long n, double *partials,
                               int nparts, hipStream_t s);
ClarityResult_t k_max_partials(const float *a, long long n, float *partials,
                               int nparts, hipStream_t s);
// out[0] = (float)sum(partials); if num != nullptr also
// scale_out[0] = num[0] / out[0]  (float division).
ClarityResult_t k_finish_sum(const double *partials, int nparts, float *out,
                             const float *num, float *scale_out,
                             hipStream_t s);
ClarityResult_t k_finish_max(const float *partials, int nparts, float *out,
                             hipStream_t s);
// h *= 1 / (float)(1.0 / sum) from float64 partials (blind PSF renorm)
ClarityResult_t k_normalize_fp64(float *h, long long n, double *partials,
                                 int nparts, hipStream_t s);
// float32 normalise h /= sum with float sum of float64 partials
ClarityResult_t k_normalize_f32sum(float *h, long long n, double *partials,
                                   int nparts, float *tmp, hipStream_t s);
ClarityResult_t k_idiv_partials(const float *g, const float *ghat, long long n,
                                double *partials, int nparts, hipStream_t s);
ClarityResult_t k_pad_shift(float *dst, Dims dd, const float *src, Dims sd,
                            const int shift[3], float fill, hipStream_t s);
ClarityResult_t k_clip(float *dst, Dims dd, const float *src, Dims sd,
                       hipStream_t s);
ClarityResult_t k_fill(float *dst, long long n, float v, hipStream_t s);

constexpr int kParts = 1024;  // float64 partials per reduction

}  // namespace clarity
