// Generic mixed-radix Stockham FFT passes for any axis length N <= 4096
// whose prime factors are <= 64 (other lengths: fft_large.hip).  One workgroup transforms a tile of TP
// pencils held interleaved in LDS (element j of pencil q at j*TP + q), so
// that loads and stores of pencils that are adjacent in memory coalesce.
//
// Stage with radix R and span p (product of earlier radices): butterfly i
// reads x[i + r*N/R], multiplies by w_{pR}^{r*(i mod p)}, takes a length-R
// DFT and writes y[(i/p)*p*R + (i mod p) + r*p]; after the last stage y is
// the DFT in natural order.
//
// These passes serve every shape; the power-of-two Richardson-Lucy pipeline
// has its own register-resident kernels.
#include "common.hpp"
#include "complex_ops.hpp"

#include <cmath>

namespace clarity {

namespace {

constexpr int kThreads = 256;

struct GenericArgs {
  int N, TP, nst, sign;
  int radix[kMaxStages];
  const float2 *tw;
  PencilGeom g;
};

__device__ __forceinline__ float2 twiddle(const float2 *tw, int idx, int sign) {
  float2 w = tw[idx];
  if (sign > 0) w.y = -w.y;
  return w;
}

template <int R>
__device__ __forceinline__ void small_dft(float2 (&u)[R], int sign,
                                          const float2 *tw, int N) {
  if constexpr (R == 2) {
    dft2(u[0], u[1]);
  } else if constexpr (R == 4) {
    dft4(u, sign);
  } else if constexpr (R == 8) {
    dft8(u, sign);
  } else {
    // direct DFT with w_R^{qs} = w_N^{(qs mod R) N/R}
    float2 v[R];
    const int m = N / R;
#pragma unroll
    for (int s = 0; s < R; s++) {
      float2 acc = u[0];
#pragma unroll
      for (int q = 1; q < R; q++) {
        float2 w = twiddle(tw, ((q * s) % R) * m, sign);
        acc = cadd(acc, cmul(u[q], w));
      }
      v[s] = acc;
    }
#pragma unroll
    for (int s = 0; s < R; s++) u[s] = v[s];
  }
}

template <int R>
__device__ __forceinline__ void stage_fixed(const float2 *src, float2 *dst,
                                            int N, int TP, int p, int sign,
                                            const float2 *tw) {
  const int m = N / R;
  const int tstep = N / (p * R);
  const int nb = TP * m;
  for (int bf = threadIdx.x; bf < nb; bf += blockDim.x) {
    const int q = bf % TP;
    const int i = bf / TP;
    const int k = i % p;
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; r++) u[r] = src[(i + r * m) * TP + q];
    if (p > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) u[r] = cmul(u[r], twiddle(tw, r * k * tstep, sign));
    }
    small_dft<R>(u, sign, tw, N);
    const int base = (i / p) * p * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) dst[(base + r * p) * TP + q] = u[r];
  }
}

// Radices above 8 that are not 16: odd primes up to 64, done with a direct
// DFT that reads inputs from LDS (no register array of runtime size).
__device__ void stage_prime(const float2 *src, float2 *dst, int N, int TP,
                            int p, int R, int sign, const float2 *tw) {
  const int m = N / R;
  const int tstep = N / (p * R);
  const int nb = TP * m * R;  // one output element per work item
  for (int w = threadIdx.x; w < nb; w += blockDim.x) {
    const int q = w % TP;
    const int rest = w / TP;
    const int s = rest % R;  // output index within the butterfly
    const int i = rest / R;
    const int k = i % p;
    float2 acc = make_float2(0.f, 0.f);
    for (int r = 0; r < R; r++) {
      float2 u = src[(i + r * m) * TP + q];
      if (p > 1 && r > 0) u = cmul(u, twiddle(tw, r * k * tstep, sign));
      float2 wv = twiddle(tw, ((r * s) % R) * m, sign);
      acc = cadd(acc, cmul(u, wv));
    }
    dst[((i / p) * p * R + k + s * p) * TP + q] = acc;
  }
}

// Runs all stages; returns the LDS buffer holding the result.
__device__ float2 *run_stages(float2 *A, float2 *B, const GenericArgs &a) {
  int p = 1;
  for (int s = 0; s < a.nst; s++) {
    const int R = a.radix[s];
    switch (R) {
      case 2: stage_fixed<2>(A, B, a.N, a.TP, p, a.sign, a.tw); break;
      case 3: stage_fixed<3>(A, B, a.N, a.TP, p, a.sign, a.tw); break;
      case 4: stage_fixed<4>(A, B, a.N, a.TP, p, a.sign, a.tw); break;
      case 5: stage_fixed<5>(A, B, a.N, a.TP, p, a.sign, a.tw); break;
      case 7: stage_fixed<7>(A, B, a.N, a.TP, p, a.sign, a.tw); break;
      case 8: stage_fixed<8>(A, B, a.N, a.TP, p, a.sign, a.tw); break;
      default: stage_prime(A, B, a.N, a.TP, p, R, a.sign, a.tw); break;
    }
    __syncthreads();
    float2 *t = A;
    A = B;
    B = t;
    p *= R;
  }
  return A;
}

struct Tile {
  long long a, b0;
  int valid;
};

__device__ __forceinline__ Tile tile_of(const PencilGeom &g, int TP) {
  const long long ntb = (g.nB + TP - 1) / TP;
  Tile t;
  t.a = (long long)blockIdx.x / ntb;
  t.b0 = ((long long)blockIdx.x % ntb) * TP;
  long long v = g.nB - t.b0;
  t.valid = v < TP ? (int)v : TP;
  return t;
}

// Element-to-thread map for a tile of TP pencils of length L: when the
// pencil elements are contiguous in memory (and the tile index is not),
// consecutive threads walk along a pencil; otherwise across pencils.
__device__ __forceinline__ void tile_index(int idx, int L, int TP,
                                           bool along_pencil, int &q, int &j) {
  if (along_pencil) {
    j = idx % L;
    q = idx / L;
  } else {
    q = idx % TP;
    j = idx / TP;
  }
}

__global__ void __launch_bounds__(kThreads)
    k_c2c_generic(const float2 *__restrict__ in, float2 *out, GenericArgs a) {
  extern __shared__ float2 smem[];
  float2 *A = smem;
  float2 *B = smem + a.TP * a.N;
  const PencilGeom &g = a.g;
  const Tile t = tile_of(g, a.TP);
  const bool in_along = (g.ie == 1 && g.ib != 1);
  const int nin = a.TP * a.N;
  for (int idx = threadIdx.x; idx < nin; idx += blockDim.x) {
    int q, j;
    tile_index(idx, a.N, a.TP, in_along, q, j);
    float2 v = make_float2(0.f, 0.f);
    if (q < t.valid && j < g.Lin)
      v = in[t.a * g.ia + (t.b0 + q) * g.ib + (long long)j * g.ie];
    A[j * a.TP + q] = v;
  }
  __syncthreads();
  float2 *res = run_stages(A, B, a);
  const bool out_along = (g.oe == 1 && g.ob != 1);
  const int nout = a.TP * g.Lout;
  for (int idx = threadIdx.x; idx < nout; idx += blockDim.x) {
    int q, j;
    tile_index(idx, g.Lout, a.TP, out_along, q, j);
    if (q < t.valid)
      out[t.a * g.oa + (t.b0 + q) * g.ob + (long long)j * g.oe] = res[j * a.TP + q];
  }
}

__global__ void __launch_bounds__(kThreads)
    k_r2c_generic(const float *__restrict__ in, float2 *out, GenericArgs a) {
  extern __shared__ float2 smem[];
  float2 *A = smem;
  float2 *B = smem + a.TP * a.N;
  const PencilGeom &g = a.g;
  const Tile t = tile_of(g, a.TP);
  const bool in_along = (g.ie == 1 && g.ib != 1);
  const int nin = a.TP * a.N;
  for (int idx = threadIdx.x; idx < nin; idx += blockDim.x) {
    int q, j;
    tile_index(idx, a.N, a.TP, in_along, q, j);
    float v = 0.f;
    if (q < t.valid && j < g.Lin)
      v = in[t.a * g.ia + (t.b0 + q) * g.ib + (long long)j * g.ie];
    A[j * a.TP + q] = make_float2(v, 0.f);
  }
  __syncthreads();
  float2 *res = run_stages(A, B, a);
  const int H = a.N / 2 + 1;
  const int L = g.Lout < H ? g.Lout : H;
  const bool out_along = (g.oe == 1 && g.ob != 1);
  const int nout = a.TP * L;
  for (int idx = threadIdx.x; idx < nout; idx += blockDim.x) {
    int q, j;
    tile_index(idx, L, a.TP, out_along, q, j);
    if (q < t.valid)
      out[t.a * g.oa + (t.b0 + q) * g.ob + (long long)j * g.oe] = res[j * a.TP + q];
  }
}

__global__ void __launch_bounds__(kThreads)
    k_c2r_generic(const float2 *__restrict__ in, float *out, GenericArgs a) {
  extern __shared__ float2 smem[];
  float2 *A = smem;
  float2 *B = smem + a.TP * a.N;
  const PencilGeom &g = a.g;
  const Tile t = tile_of(g, a.TP);
  const int N = a.N;
  const int H = N / 2 + 1;
  const bool in_along = (g.ie == 1 && g.ib != 1);
  const int nin = a.TP * H;
  for (int idx = threadIdx.x; idx < nin; idx += blockDim.x) {
    int q, j;
    tile_index(idx, H, a.TP, in_along, q, j);
    float2 v = make_float2(0.f, 0.f);
    if (q < t.valid) v = in[t.a * g.ia + (t.b0 + q) * g.ib + (long long)j * g.ie];
    // The imaginary parts of bin 0 and (for even N) bin N/2 are ignored.
    if (j == 0 || 2 * j == N) v.y = 0.f;
    A[j * a.TP + q] = v;
    if (j > 0 && 2 * j != N) A[(N - j) * a.TP + q] = make_float2(v.x, -v.y);
  }
  __syncthreads();
  float2 *res = run_stages(A, B, a);
  const bool out_along = (g.oe == 1 && g.ob != 1);
  const int nout = a.TP * g.Lout;
  for (int idx = threadIdx.x; idx < nout; idx += blockDim.x) {
    int q, j;
    tile_index(idx, g.Lout, a.TP, out_along, q, j);
    if (q < t.valid)
      out[t.a * g.oa + (t.b0 + q) * g.ob + (long long)j * g.oe] = res[j * a.TP + q].x;
  }
}

int pick_tp(int N) {
  int tp = kMaxGenericN / N;
  if (tp > 16) tp = 16;
  if (tp < 1) tp = 1;
  return tp;
}

ClarityResult_t make_args(const AxisPlan &ap, const PencilGeom &g, int sign,
                          GenericArgs *a, size_t *lds, long long *grid) {
  if (ap.N <= 0 || ap.N > kMaxGenericN || !ap.tw) return CLARITY_FFT_FAILED;
  a->N = ap.N;
  a->TP = pick_tp(ap.N);
  a->nst = ap.nst;
  a->sign = sign;
  for (int i = 0; i < kMaxStages; i++) a->radix[i] = ap.radix[i];
  a->tw = ap.tw;
  a->g = g;
  *lds = sizeof(float2) * 2 * (size_t)a->TP * ap.N;
  *grid = g.nA * ((g.nB + a->TP - 1) / a->TP);
  if (*grid <= 0) return CLARITY_SUCCESS;
  if (*grid > 0x7fffffffLL) return CLARITY_FFT_FAILED;
  return CLARITY_SUCCESS;
}

}  // namespace

bool fft_length_direct(int n) {
  AxisPlan ap;
  return axis_factor(n, &ap) == CLARITY_SUCCESS;
}

// Factorises n into radices 8, 4, 2 (powers of two first), then 3, 5, 7 and
// other odd primes up to 64.
ClarityResult_t axis_factor(int n, AxisPlan *out) {
  if (n <= 0 || n > kMaxGenericN) return CLARITY_FFT_FAILED;
  AxisPlan ap;
  ap.N = n;
  int m = n;
  int e2 = 0;
  while (m % 2 == 0) { m /= 2; e2++; }
  while (e2 >= 3) { ap.radix[ap.nst++] = 8; e2 -= 3; }
  if (e2 == 2) ap.radix[ap.nst++] = 4;
  if (e2 == 1) ap.radix[ap.nst++] = 2;
  for (int p = 3; m > 1; p += 2) {
    while (m % p == 0) {
      if (p > 64 || ap.nst >= kMaxStages) return CLARITY_FFT_FAILED;
      ap.radix[ap.nst++] = p;
      m /= p;
    }
  }
  *out = ap;
  return CLARITY_SUCCESS;
}

namespace {

__global__ void twiddle_kernel(float2 *tw, int n) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const double a = -2.0 * M_PI * (double)j / (double)n;
    tw[j] = make_float2((float)cos(a), (float)sin(a));
  }
}

}  // namespace

ClarityResult_t k_twiddles(float2 *tw, int n, hipStream_t s) {
  if (n <= 0) return CLARITY_INVALID_ARGUMENT;
  const int blocks = (n + 255) / 256;
  hipLaunchKernelGGL(twiddle_kernel, dim3(blocks), dim3(256), 0, s, tw, n);
  return launch_check();
}

// axis_factor, then builds (or reuses) the w_n table on the device.
ClarityResult_t axis_plan(int n, AxisPlan *out) {
  AxisPlan ap;
  CL_TRY(axis_factor(n, &ap));
  Runtime &r = rt();
  auto it = r.twiddles.find(n);
  if (it == r.twiddles.end()) {
    std::vector<float2> h(n);
    for (int j = 0; j < n; j++) {
      // exact reductions keep the table symmetric: angle in [0, 2 pi)
      double a = -2.0 * M_PI * (double)j / (double)n;
      h[j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    void *d = nullptr;
    CL_TRY(dev_alloc(&d, sizeof(float2) * n));
    CL_HIP(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    r.twiddles[n] = static_cast<float2 *>(d);
    ap.tw = static_cast<float2 *>(d);
  } else {
    ap.tw = it->second;
  }
  *out = ap;
  return CLARITY_SUCCESS;
}

ClarityResult_t fft_c2c_pass(const float2 *in, float2 *out, const AxisPlan &ap,
                             const PencilGeom &g, int sign, hipStream_t s) {
  GenericArgs a;
  size_t lds;
  long long grid;
  CL_TRY(make_args(ap, g, sign, &a, &lds, &grid));
  if (grid == 0) return CLARITY_SUCCESS;
  hipLaunchKernelGGL(k_c2c_generic, dim3((unsigned)grid), dim3(kThreads), lds,
                     s, in, out, a);
  return launch_check();
}

ClarityResult_t fft_r2c_rows(const float *in, float2 *out, const AxisPlan &ap,
                             const PencilGeom &g, hipStream_t s) {
  GenericArgs a;
  size_t lds;
  long long grid;
  CL_TRY(make_args(ap, g, -1, &a, &lds, &grid));
  if (grid == 0) return CLARITY_SUCCESS;
  hipLaunchKernelGGL(k_r2c_generic, dim3((unsigned)grid), dim3(kThreads), lds,
                     s, in, out, a);
  return launch_check();
}

ClarityResult_t fft_c2r_rows(const float2 *in, float *out, const AxisPlan &ap,
                             const PencilGeom &g, hipStream_t s) {
  GenericArgs a;
  size_t lds;
  long long grid;
  CL_TRY(make_args(ap, g, +1, &a, &lds, &grid));
  if (grid == 0) return CLARITY_SUCCESS;
  hipLaunchKernelGGL(k_c2r_generic, dim3((unsigned)grid), dim3(kThreads), lds,
                     s, in, out, a);
  return launch_check();
}

namespace {

// Pencils of the y and z passes over a half spectrum (z, y, kx).
PencilGeom y_pencils(int ny, int nz, long long hx) {
  PencilGeom g;  // pencils (z, kx) along y
  g.nA = nz; g.nB = hx;
  g.ia = g.oa = (long long)ny * hx; g.ib = g.ob = 1; g.ie = g.oe = hx;
  g.Lin = ny; g.Lout = ny;
  return g;
}

PencilGeom z_pencils(int ny, int nz, long long hx) {
  PencilGeom g;  // pencils (y, kx) along z: one outer block
  g.nA = 1; g.nB = (long long)ny * hx;
  g.ia = g.oa = 0; g.ib = g.ob = 1; g.ie = g.oe = (long long)ny * hx;
  g.Lin = nz; g.Lout = nz;
  return g;
}

// A complex pass along y or z over the spectrum in place: the LDS passes
// for their lengths, fft_axis_large for the others.
ClarityResult_t spectrum_pass(float2 *spec, int n, const PencilGeom &g, int sign,
                              hipStream_t s) {
  if (fft_length_direct(n)) {
    AxisPlan ap;
    CL_TRY(axis_plan(n, &ap));
    return fft_c2c_pass(spec, spec, ap, g, sign, s);
  }
  return fft_axis_large(spec, spec, g, n, kGatherComplex, sign, n, false, s);
}

}  // namespace

// FFTW layout: real (z, y, x) -> spectrum (z, y, kx), hx = nx/2+1.  Every
// axis length up to kMaxFftLength.
ClarityResult_t fft3d_r2c(int nx, int ny, int nz, const float *in, float2 *out,
                          hipStream_t s) {
  if (!fft_length_supported(nx) || !fft_length_supported(ny) || !fft_length_supported(nz))
    return CLARITY_FFT_FAILED;
  if (fft_length_direct(nx) && fft_length_direct(ny) && fft_length_direct(nz)) {
    AxisPlan px, py, pz;
    CL_TRY(axis_plan(nx, &px));
    CL_TRY(axis_plan(ny, &py));
    CL_TRY(axis_plan(nz, &pz));
    return fft3d_r2c_planned(px, py, pz, in, out, s);
  }
  const long long hx = nx / 2 + 1;
  PencilGeom gx;  // rows (z, y): a = z, b = y
  gx.nA = nz; gx.nB = ny;
  gx.ia = (long long)ny * nx; gx.ib = nx; gx.ie = 1;
  gx.oa = (long long)ny * hx; gx.ob = hx; gx.oe = 1;
  gx.Lin = nx; gx.Lout = (int)hx;
  if (fft_length_direct(nx)) {
    AxisPlan px;
    CL_TRY(axis_plan(nx, &px));
    CL_TRY(fft_r2c_rows(in, out, px, gx, s));
  } else {
    CL_TRY(fft_axis_large(in, out, gx, nx, kGatherReal, -1, (int)hx, false, s));
  }
  CL_TRY(spectrum_pass(out, ny, y_pencils(ny, nz, hx), -1, s));
  return spectrum_pass(out, nz, z_pencils(ny, nz, hx), -1, s);
}

ClarityResult_t fft3d_r2c_planned(const AxisPlan &px, const AxisPlan &py,
                                  const AxisPlan &pz, const float *in, float2 *out,
                                  hipStream_t s) {
  const int nx = px.N, ny = py.N, nz = pz.N;
  const long long hx = nx / 2 + 1;
  PencilGeom gx;  // rows (z, y): a = z, b = y
  gx.nA = nz; gx.nB = ny;
  gx.ia = (long long)ny * nx; gx.ib = nx; gx.ie = 1;
  gx.oa = (long long)ny * hx; gx.ob = hx; gx.oe = 1;
  gx.Lin = nx; gx.Lout = (int)hx;
  CL_TRY(fft_r2c_rows(in, out, px, gx, s));
  PencilGeom gy;  // pencils (z, kx) along y
  gy.nA = nz; gy.nB = hx;
  gy.ia = gy.oa = (long long)ny * hx; gy.ib = gy.ob = 1; gy.ie = gy.oe = hx;
  gy.Lin = ny; gy.Lout = ny;
  CL_TRY(fft_c2c_pass(out, out, py, gy, -1, s));
  PencilGeom gz;  // pencils (y, kx) along z: one outer block
  gz.nA = 1; gz.nB = (long long)ny * hx;
  gz.ia = gz.oa = 0; gz.ib = gz.ob = 1; gz.ie = gz.oe = (long long)ny * hx;
  gz.Lin = nz; gz.Lout = nz;
  return fft_c2c_pass(out, out, pz, gz, -1, s);
}

ClarityResult_t fft3d_c2r(int nx, int ny, int nz, float2 *spec, float *out,
                          hipStream_t s) {
  if (!fft_length_supported(nx) || !fft_length_supported(ny) || !fft_length_supported(nz))
    return CLARITY_FFT_FAILED;
  if (fft_length_direct(nx) && fft_length_direct(ny) && fft_length_direct(nz)) {
    AxisPlan px, py, pz;
    CL_TRY(axis_plan(nx, &px));
    CL_TRY(axis_plan(ny, &py));
    CL_TRY(axis_plan(nz, &pz));
    return fft3d_c2r_planned(px, py, pz, spec, out, s);
  }
  const long long hx = nx / 2 + 1;
  CL_TRY(spectrum_pass(spec, nz, z_pencils(ny, nz, hx), +1, s));
  CL_TRY(spectrum_pass(spec, ny, y_pencils(ny, nz, hx), +1, s));
  PencilGeom gx;
  gx.nA = nz; gx.nB = ny;
  gx.ia = (long long)ny * hx; gx.ib = hx; gx.ie = 1;
  gx.oa = (long long)ny * nx; gx.ob = nx; gx.oe = 1;
  gx.Lin = (int)hx; gx.Lout = nx;
  if (fft_length_direct(nx)) {
    AxisPlan px;
    CL_TRY(axis_plan(nx, &px));
    return fft_c2r_rows(spec, out, px, gx, s);
  }
  return fft_axis_large(spec, out, gx, nx, kGatherHalf, +1, nx, true, s);
}

ClarityResult_t fft3d_c2r_planned(const AxisPlan &px, const AxisPlan &py,
                                  const AxisPlan &pz, float2 *spec, float *out,
                                  hipStream_t s) {
  const int nx = px.N, ny = py.N, nz = pz.N;
  const long long hx = nx / 2 + 1;
  PencilGeom gz;
  gz.nA = 1; gz.nB = (long long)ny * hx;
  gz.ia = gz.oa = 0; gz.ib = gz.ob = 1; gz.ie = gz.oe = (long long)ny * hx;
  gz.Lin = nz; gz.Lout = nz;
  CL_TRY(fft_c2c_pass(spec, spec, pz, gz, +1, s));
  PencilGeom gy;
  gy.nA = nz; gy.nB = hx;
  gy.ia = gy.oa = (long long)ny * hx; gy.ib = gy.ob = 1; gy.ie = gy.oe = hx;
  gy.Lin = ny; gy.Lout = ny;
  CL_TRY(fft_c2c_pass(spec, spec, py, gy, +1, s));
  PencilGeom gx;
  gx.nA = nz; gx.nB = ny;
  gx.ia = (long long)ny * hx; gx.ib = hx; gx.ie = 1;
  gx.oa = (long long)ny * nx; gx.ob = nx; gx.oe = 1;
  gx.Lin = (int)hx; gx.Lout = nx;
  return fft_c2r_rows(spec, out, px, gx, s);
}

}  // namespace clarity
