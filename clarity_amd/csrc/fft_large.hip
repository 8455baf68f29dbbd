// Axes the LDS-resident generic passes cannot take in one tile: lengths above
// 4096 and lengths with a prime factor above 64.  FFTW, which the
// reference's FFT calls (src/FFT.cxx:46-146), plans every length; these
// transforms make the library do the same.
//
// The pencils of such an axis are gathered into an element-major scratch
// layout T[m * P + p] (element m of pencil p; adjacent pencils adjacent in
// memory, so gathers, scatters and every pass coalesce), transformed there
// as a batch, and scattered back:
//
//  * four-step, for lengths L = L1 * L2 with both factors direct lengths of
//    the generic passes (L1, L2 <= 4096, prime factors <= 64): with
//    j = L2 j1 + j2 and k = k1 + L1 k2,
//      X[k1 + L1 k2] = sum_j2 w_L^(j2 k1) w_L2^(j2 k2) sum_j1 x[L2 j1 + j2] w_L1^(j1 k1),
//    i.e. length-L1 transforms in place, a twiddle multiplication, and
//    length-L2 transforms written out of place in natural order;
//  * Bluestein (chirp-z), for every other length n: with c_j = w^(j^2/2),
//    w = exp(sign 2 pi i / n), X_k = c_k sum_j (x_j c_j) conj(c_(k-j)), a
//    linear convolution computed as a cyclic one of power-of-two length
//    M >= 2n - 1 (forward transform, product with the transform of the
//    conjugate chirp, inverse transform; the M-point transforms direct or
//    four-step).
//
// Tables (chirp, the chirp's transform, four-step twiddles) are computed on
// the host in float64 and cached per length like the generic twiddles.
#include "common.hpp"
#include "complex_ops.hpp"

#include <cmath>
#include <complex>
#include <vector>

namespace clarity {

namespace {

constexpr int kBlock = 256;

unsigned grid_for(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > 1048576) b = 1048576;
  return (unsigned)b;
}

#define GRID_STRIDE(i, n)                                                    \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (long long)gridDim.x * blockDim.x)


// T[m * P + p] = element m of pencil p (times chirp[m] when chirp is given)
// for m < n, and 0 for n <= m < L.  Pencil p = a * nB + b has element m at
// a * ia + b * ib + m * ie of `in` (float or float2 units).
__global__ void gather_kernel(float2 *__restrict__ T, const void *__restrict__ in,
                              PencilGeom g, int n, long long L, int kind,
                              const float2 *__restrict__ chirp) {
  const long long P = g.nA * g.nB;
  GRID_STRIDE(i, L * P) {
    const long long m = i / P, p = i % P;
    const long long a = p / g.nB, b = p % g.nB;
    float2 v = make_float2(0.f, 0.f);
    if (m < n) {
      if (kind == kGatherReal) {
        v.x = static_cast<const float *>(in)[a * g.ia + b * g.ib + m * g.ie];
      } else if (kind == kGatherComplex) {
        v = static_cast<const float2 *>(in)[a * g.ia + b * g.ib + m * g.ie];
      } else {
        const long long h = n / 2;
        const long long src = m <= h ? m : n - m;
        v = static_cast<const float2 *>(in)[a * g.ia + b * g.ib + src * g.ie];
        if (m == 0 || 2 * m == n) v.y = 0.f;
        else if (m > h) v.y = -v.y;
      }
      if (chirp) v = cmul(v, chirp[m]);
    }
    T[i] = v;
  }
}

// Element k < count of pencil p of `out` = R[k * P + p] (times chirp[k]
// when given): complex, or its real part when real_out.
__global__ void scatter_kernel(void *out, const float2 *__restrict__ R, PencilGeom g,
                               int count, int real_out, const float2 *__restrict__ chirp) {
  const long long P = g.nA * g.nB;
  GRID_STRIDE(i, (long long)count * P) {
    const long long k = i / P, p = i % P;
    const long long a = p / g.nB, b = p % g.nB;
    float2 v = R[i];
    if (chirp) v = cmul(v, chirp[k]);
    const long long o = a * g.oa + b * g.ob + k * g.oe;
    if (real_out)
      static_cast<float *>(out)[o] = v.x;
    else
      static_cast<float2 *>(out)[o] = v;
  }
}

// T[m * P + p] *= f[m]
__global__ void scale_rows_kernel(float2 *T, long long P, long long L,
                                  const float2 *__restrict__ f) {
  GRID_STRIDE(i, L * P) T[i] = cmul(T[i], f[i / P]);
}

// Four-step twiddle: T[(L2 k1 + j2) * P + p] *= w_L^(j2 k1), tw[j] =
// exp(-2 pi i j / L), conjugated for sign > 0.
__global__ void fourstep_twiddle_kernel(float2 *T, long long P, int L1, int L2, int sign,
                                        const float2 *__restrict__ tw) {
  const long long L = (long long)L1 * L2;
  GRID_STRIDE(i, L * P) {
    const long long m = i / P;
    const long long k1 = m / L2, j2 = m % L2;
    float2 w = tw[(j2 * k1) % L];
    if (sign > 0) w.y = -w.y;
    T[i] = cmul(T[i], w);
  }
}

bool direct_length(int n) { return fft_length_direct(n); }

// L = L1 * L2 with both factors direct lengths, the most balanced such
// split; false if there is none.
bool fourstep_split(long long L, int *L1, int *L2) {
  if (L > (long long)kMaxGenericN * kMaxGenericN) return false;
  int best = 0;
  long long bestd = -1;
  for (int a = 2; a <= kMaxGenericN; a++) {
    if (L % a) continue;
    const long long b = L / a;
    if (b < 2 || b > kMaxGenericN) continue;
    if (!direct_length(a) || !direct_length((int)b)) continue;
    const long long d = a > b ? a - b : b - a;
    if (bestd < 0 || d < bestd) {
      best = a;
      bestd = d;
    }
  }
  if (!best) return false;
  *L1 = best;
  *L2 = (int)(L / best);
  return true;
}

// Device tables cached in the runtime (freed at the last UnRegister).
template <typename Fill>
ClarityResult_t table(long long key, long long count, Fill fill, const float2 **out) {
  Runtime &r = rt();
  auto it = r.tables.find(key);
  if (it != r.tables.end()) {
    *out = it->second;
    return CLARITY_SUCCESS;
  }
  std::vector<float2> h((size_t)count);
  fill(h);
  void *d = nullptr;
  CL_TRY(dev_alloc(&d, sizeof(float2) * (size_t)count));
  CL_HIP(hipMemcpy(d, h.data(), sizeof(float2) * (size_t)count, hipMemcpyHostToDevice));
  r.tables[key] = static_cast<float2 *>(d);
  *out = static_cast<float2 *>(d);
  return CLARITY_SUCCESS;
}

// keys: kind in the top bits, then sign, then the length
long long table_key(int kind, int sign, long long n) {
  return ((long long)kind << 56) | ((long long)(sign > 0) << 48) | n;
}

// exp(-2 pi i j / L), j < L
ClarityResult_t fourstep_table(long long L, const float2 **out) {
  return table(table_key(1, -1, L), L, [L](std::vector<float2> &h) {
    for (long long j = 0; j < L; j++) {
      const double a = -2.0 * M_PI * (double)j / (double)L;
      h[(size_t)j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
  }, out);
}

// c_j = exp(sign pi i j^2 / n), j < n (j^2 reduced mod 2n exactly)
ClarityResult_t chirp_table(int n, int sign, const float2 **out) {
  return table(table_key(2, sign, n), n, [n, sign](std::vector<float2> &h) {
    for (long long j = 0; j < n; j++) {
      const long long q = (j * j) % (2LL * n);
      const double a = (double)sign * M_PI * (double)q / (double)n;
      h[(size_t)j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
  }, out);
}

// In-place iterative radix-2 forward DFT in float64 (M a power of two),
// twiddles from one table of exp(-2 pi i k / M), k < M/2.
void host_fft_pow2(std::vector<std::complex<double>> &x) {
  const size_t M = x.size();
  if (M < 2) return;
  std::vector<std::complex<double>> w(M / 2);
  for (size_t k = 0; k < M / 2; k++) {
    const double a = -2.0 * M_PI * (double)k / (double)M;
    w[k] = std::complex<double>(std::cos(a), std::sin(a));
  }
  for (size_t i = 1, j = 0; i < M; i++) {
    size_t bit = M >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(x[i], x[j]);
  }
  for (size_t len = 2; len <= M; len <<= 1) {
    const size_t step = M / len;
    for (size_t i = 0; i < M; i += len) {
      for (size_t k = 0; k < len / 2; k++) {
        const std::complex<double> u = x[i + k], v = x[i + k + len / 2] * w[k * step];
        x[i + k] = u + v;
        x[i + k + len / 2] = u - v;
      }
    }
  }
}

// The forward M-point transform of the cyclic sequence d[m] = conj(c_m)
// for m < n, d[M - m] = conj(c_m) for 0 < m < n, 0 between; divided by M
// (the inverse transform's normalisation folded in).
ClarityResult_t chirp_filter_table(int n, int sign, long long M, const float2 **out) {
  return table(table_key(3, sign, ((long long)n << 24) | M), M,
               [n, sign, M](std::vector<float2> &h) {
    std::vector<std::complex<double>> d((size_t)M, 0.0);
    for (long long m = 0; m < n; m++) {
      const long long q = (m * m) % (2LL * n);
      const double a = -(double)sign * M_PI * (double)q / (double)n;
      const std::complex<double> c(std::cos(a), std::sin(a));
      d[(size_t)m] = c;
      if (m > 0) d[(size_t)(M - m)] = c;
    }
    host_fft_pow2(d);
    for (long long m = 0; m < M; m++)
      h[(size_t)m] = make_float2((float)(d[(size_t)m].real() / (double)M),
                                 (float)(d[(size_t)m].imag() / (double)M));
  }, out);
}

// Batched transform of T[m * P + p], m < L: direct pass or four-step.  The
// result is in *res, which is T or U (U: scratch of the same size).
ClarityResult_t batch_fft(float2 *T, float2 *U, long long P, long long L, int sign,
                          hipStream_t s, float2 **res) {
  if (L <= kMaxGenericN && direct_length((int)L)) {
    AxisPlan ap;
    CL_TRY(axis_plan((int)L, &ap));
    PencilGeom g;
    g.nA = 1; g.nB = P;
    g.ia = g.oa = 0; g.ib = g.ob = 1; g.ie = g.oe = P;
    g.Lin = g.Lout = (int)L;
    CL_TRY(fft_c2c_pass(T, T, ap, g, sign, s));
    *res = T;
    return CLARITY_SUCCESS;
  }
  int L1 = 0, L2 = 0;
  if (!fourstep_split(L, &L1, &L2)) return CLARITY_FFT_FAILED;
  AxisPlan a1, a2;
  CL_TRY(axis_plan(L1, &a1));
  CL_TRY(axis_plan(L2, &a2));
  const float2 *tw = nullptr;
  CL_TRY(fourstep_table(L, &tw));
  // length-L1 transforms over j1 (stride L2 P) of pencils (j2, p), in place
  PencilGeom g1;
  g1.nA = 1; g1.nB = (long long)L2 * P;
  g1.ia = g1.oa = 0; g1.ib = g1.ob = 1; g1.ie = g1.oe = (long long)L2 * P;
  g1.Lin = g1.Lout = L1;
  CL_TRY(fft_c2c_pass(T, T, a1, g1, sign, s));
  hipLaunchKernelGGL(fourstep_twiddle_kernel, dim3(grid_for(L * P)), dim3(kBlock), 0, s, T,
                     P, L1, L2, sign, tw);
  CL_TRY(launch_check());
  // length-L2 transforms over j2 (stride P) of pencils (k1, p), written to
  // U at element k1 + L1 k2
  PencilGeom g2;
  g2.nA = L1; g2.nB = P;
  g2.ia = (long long)L2 * P; g2.ib = 1; g2.ie = P;
  g2.oa = P; g2.ob = 1; g2.oe = (long long)L1 * P;
  g2.Lin = g2.Lout = L2;
  CL_TRY(fft_c2c_pass(T, U, a2, g2, sign, s));
  *res = U;
  return CLARITY_SUCCESS;
}

long long pow2_at_least(long long v) {
  long long m = 1;
  while (m < v) m <<= 1;
  return m;
}

}  // namespace

bool fft_length_supported(long long n) {
  // direct, four-step, or Bluestein with a power-of-two inner length M <=
  // 2^24 = 4096^2 (itself direct or four-step)
  return n > 0 && n <= kMaxFftLength;
}

// One axis of length n over the pencils of `g` (P = nA nB pencils, input
// at a ia + b ib + m ie, output at a oa + b ob + k oe): gathers (kind),
// transforms with `sign`, and scatters the first `count` outputs (complex,
// or real parts when real_out).  in may equal out.
ClarityResult_t fft_axis_large(const void *in, void *out, const PencilGeom &g, int n,
                               int kind, int sign, int count, bool real_out, hipStream_t s) {
  if (!fft_length_supported(n)) return CLARITY_FFT_FAILED;
  const long long P = g.nA * g.nB;
  if (P <= 0) return CLARITY_SUCCESS;
  int L1, L2;
  const bool bluestein = !(n <= kMaxGenericN && direct_length(n)) && !fourstep_split(n, &L1, &L2);
  const long long L = bluestein ? pow2_at_least(2LL * n - 1) : n;
  DevBuf Tb, Ub;
  CL_TRY(Tb.alloc(sizeof(float2) * (size_t)(L * P)));
  if (L > kMaxGenericN) CL_TRY(Ub.alloc(sizeof(float2) * (size_t)(L * P)));
  float2 *T = Tb.c(), *U = Ub.c();
  const float2 *chirp = nullptr, *filt = nullptr;
  if (bluestein) {
    CL_TRY(chirp_table(n, sign, &chirp));
    CL_TRY(chirp_filter_table(n, sign, L, &filt));
  }
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(L * P)), dim3(kBlock), 0, s, T, in, g, n,
                     L, kind, chirp);
  CL_TRY(launch_check());
  float2 *R = nullptr;
  if (!bluestein) {
    CL_TRY(batch_fft(T, U, P, L, sign, s, &R));
  } else {
    CL_TRY(batch_fft(T, U, P, L, -1, s, &R));
    hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_for(L * P)), dim3(kBlock), 0, s, R, P, L,
                       filt);
    CL_TRY(launch_check());
    float2 *R2 = nullptr;
    CL_TRY(batch_fft(R, R == T ? U : T, P, L, +1, s, &R2));
    R = R2;
  }
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for((long long)count * P)), dim3(kBlock), 0, s,
                     out, R, g, count, real_out ? 1 : 0, chirp);
  CL_TRY(launch_check());
  // the buffers go back to the cache only once the stream is past them
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

}  // namespace clarity
