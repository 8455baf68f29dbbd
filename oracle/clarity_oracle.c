/*
 * clarity_oracle.c -- CPU restatement of Clarity's FFT deconvolution path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline timed by bench.py.  The product never links or calls this file.
 *
 * Each function names the reference function it restates (paths relative to
 * the reference tree).  Element order and float32 operation order follow the
 * reference; the FFT is either this file's own float64 mixed-radix FFT
 * (default, rounded to float32 once) or the FFTW3 float API resolved at run
 * time (see clarity_oracle.h).  Build: oracle/Makefile, with
 * -ffp-contract=off so that a*c - b*d is evaluated with two roundings as in
 * the reference's x86-64 default build.
 */
#define _GNU_SOURCE
#include "clarity_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* configuration                                                            */
/* ------------------------------------------------------------------------ */

static int g_reduce_mode = ORACLE_REDUCE_FP64;
static int g_fft_backend = ORACLE_FFT_OWN;
static int g_threads = 0; /* 0 = OpenMP default */
static double g_loop_seconds = 0.0; /* wall time of the last RL loop */
/* wall time of the last RL call's deconvolution scope: from after padding to
 * before the clip (src/MaximumLikelihoodDeconvolve.cxx:294-316,
 * TIME_DECONVOLVE: the CPU driver's buffers, the kernel transform, the
 * energy and the iterations) */
static double g_deconv_seconds = 0.0;

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double oracle_last_loop_seconds(void) { return g_loop_seconds; }
double oracle_last_deconvolve_seconds(void) { return g_deconv_seconds; }

void oracle_set_reduce_mode(int mode) { g_reduce_mode = mode; }
int oracle_get_reduce_mode(void) { return g_reduce_mode; }
int oracle_get_fft_backend(void) { return g_fft_backend; }

void oracle_set_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}

int oracle_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* FFTW3 float API, resolved at run time (src/FFT.cxx:80-86,131-137).       */
/* ------------------------------------------------------------------------ */

typedef void *fw_plan;
typedef fw_plan (*fw_r2c_3d_t)(int, int, int, float *, float *, unsigned);
typedef fw_plan (*fw_c2r_3d_t)(int, int, int, float *, float *, unsigned);
typedef void (*fw_exec_t)(fw_plan);
typedef void (*fw_destroy_t)(fw_plan);
typedef int (*fw_init_threads_t)(void);
typedef void (*fw_nthreads_t)(int);
#define FW_ESTIMATE (1U << 6)

static struct {
  int loaded;
  char name[128];
  fw_r2c_3d_t r2c;
  fw_c2r_3d_t c2r;
  fw_exec_t exec;
  fw_destroy_t destroy;
  fw_nthreads_t nthreads;
} g_fw;

static int fw_try(const char *lib, const char *threads_lib, const char *label) {
  void *h = dlopen(lib, RTLD_NOW | RTLD_GLOBAL);
  if (!h) return 0;
  void *ht = threads_lib ? dlopen(threads_lib, RTLD_NOW | RTLD_GLOBAL) : h;
  if (!ht) ht = h;
  g_fw.r2c = (fw_r2c_3d_t)dlsym(h, "fftwf_plan_dft_r2c_3d");
  g_fw.c2r = (fw_c2r_3d_t)dlsym(h, "fftwf_plan_dft_c2r_3d");
  g_fw.exec = (fw_exec_t)dlsym(h, "fftwf_execute");
  g_fw.destroy = (fw_destroy_t)dlsym(h, "fftwf_destroy_plan");
  fw_init_threads_t it = (fw_init_threads_t)dlsym(ht, "fftwf_init_threads");
  g_fw.nthreads = (fw_nthreads_t)dlsym(ht, "fftwf_plan_with_nthreads");
  if (!g_fw.r2c || !g_fw.c2r || !g_fw.exec || !g_fw.destroy) return 0;
  /* src/Clarity.cxx:49 and :93-103: init threads, plan with np threads */
  if (it) it();
  snprintf(g_fw.name, sizeof g_fw.name, "%s", label);
  g_fw.loaded = 1;
  return 1;
}

static int fw_load(void) {
  if (g_fw.loaded) return 1;
  if (fw_try("libfftw3f.so.3", "libfftw3f_threads.so.3", "FFTW3f")) return 1;
  /* MKL's FFTW3 interface; keep MKL on the GNU OpenMP runtime this file
   * already uses. */
  setenv("MKL_THREADING_LAYER", "GNU", 0);
  if (fw_try("libmkl_rt.so", NULL, "MKL-FFTW3")) return 1;
  if (fw_try("/opt/conda/lib/libmkl_rt.so", NULL, "MKL-FFTW3")) return 1;
  if (fw_try("libmkl_rt.so.1", NULL, "MKL-FFTW3")) return 1;
  if (fw_try("/opt/conda/lib/libmkl_rt.so.1", NULL, "MKL-FFTW3")) return 1;
  return 0;
}

int oracle_set_fft_backend(int backend) {
  if (backend == ORACLE_FFT_FFTW_API) {
    if (!fw_load()) return OR_FFT_FAILED;
  }
  g_fft_backend = backend;
  return OR_SUCCESS;
}

const char *oracle_fft_backend_name(void) {
  if (g_fft_backend == ORACLE_FFT_FFTW_API && g_fw.loaded) return g_fw.name;
  return "own-fp64";
}

/* ------------------------------------------------------------------------ */
/* Own float64 FFT: mixed-radix Stockham autosort.                           */
/* Stage with radix R and span p (product of earlier radices): butterfly i   */
/* reads x[i + r*N/R], multiplies by w_{pR}^{r*(i mod p)}, takes a length-R  */
/* DFT and writes y[(i/p)*p*R + (i mod p) + r*p].                            */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, nf;
  int fac[64];
  double *wr, *wi; /* w_n^j = exp(-2 pi i j / n), j < n */
} fplan;

static void fplan_init(fplan *pl, int n) {
  pl->n = n;
  pl->nf = 0;
  int m = n;
  while (m % 4 == 0) { pl->fac[pl->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { pl->fac[pl->nf++] = 2; m /= 2; }
  for (int p = 3; m > 1; p += 2) {
    while (m % p == 0) { pl->fac[pl->nf++] = p; m /= p; }
    if ((long)p * p > m && m > 1) { pl->fac[pl->nf++] = m; m = 1; }
  }
  pl->wr = (double *)malloc(sizeof(double) * n);
  pl->wi = (double *)malloc(sizeof(double) * n);
  for (int j = 0; j < n; j++) {
    double a = -2.0 * M_PI * (double)j / (double)n;
    pl->wr[j] = cos(a);
    pl->wi[j] = sin(a);
  }
}

static void fplan_free(fplan *pl) {
  free(pl->wr);
  free(pl->wi);
}

/* x: interleaved complex (2n doubles), y: scratch of the same size.  The
 * result is left in x.  sign = -1 forward, +1 inverse (unnormalised). */
static void fft_pencil(const fplan *pl, double *x, double *y, int sign) {
  const int n = pl->n;
  if (n == 1) return;
  double *src = x, *dst = y;
  int p = 1;
  double ur[64], ui[64], vr[64], vi[64];
  for (int s = 0; s < pl->nf; s++) {
    const int R = pl->fac[s];
    const int m = n / R;
    const int tstep = n / (p * R);
    for (int i = 0; i < m; i++) {
      const int k = i % p;
      if (R <= 64) {
        for (int r = 0; r < R; r++) {
          double a = src[2 * (i + r * m)], b = src[2 * (i + r * m) + 1];
          int ti = r * k * tstep;
          double c = pl->wr[ti], d = sign < 0 ? pl->wi[ti] : -pl->wi[ti];
          ur[r] = a * c - b * d;
          ui[r] = a * d + b * c;
        }
        if (R == 2) {
          vr[0] = ur[0] + ur[1]; vi[0] = ui[0] + ui[1];
          vr[1] = ur[0] - ur[1]; vi[1] = ui[0] - ui[1];
        } else if (R == 4) {
          /* X_q = sum_r u_r w^(rq), w = exp(sign 2 pi i / 4) = sign * i */
          const double s = (double)sign;
          const double ar = ur[0] + ur[2], ai = ui[0] + ui[2];
          const double br = ur[0] - ur[2], bi = ui[0] - ui[2];
          const double cr = ur[1] + ur[3], ci = ui[1] + ui[3];
          const double dr = ur[1] - ur[3], di = ui[1] - ui[3];
          vr[0] = ar + cr; vi[0] = ai + ci;
          vr[2] = ar - cr; vi[2] = ai - ci;
          /* (sign i) d = (-s di, s dr) */
          vr[1] = br - s * di; vi[1] = bi + s * dr;
          vr[3] = br + s * di; vi[3] = bi - s * dr;
        } else {
          for (int q = 0; q < R; q++) {
            double sr = 0.0, si = 0.0;
            for (int r = 0; r < R; r++) {
              int ti = (int)(((long)r * q) % R) * m;
              double c = pl->wr[ti], d = sign < 0 ? pl->wi[ti] : -pl->wi[ti];
              sr += ur[r] * c - ui[r] * d;
              si += ur[r] * d + ui[r] * c;
            }
            vr[q] = sr;
            vi[q] = si;
          }
        }
        const int base = (i / p) * p * R + k;
        for (int r = 0; r < R; r++) {
          dst[2 * (base + r * p)] = vr[r];
          dst[2 * (base + r * p) + 1] = vi[r];
        }
      } else {
        /* large prime radix: direct DFT with heap temporaries */
        double *tr = (double *)malloc(sizeof(double) * 2 * R);
        double *ti_ = tr + R;
        for (int r = 0; r < R; r++) {
          double a = src[2 * (i + r * m)], b = src[2 * (i + r * m) + 1];
          int ti = (int)(((long)r * k * tstep) % n);
          double c = pl->wr[ti], d = sign < 0 ? pl->wi[ti] : -pl->wi[ti];
          tr[r] = a * c - b * d;
          ti_[r] = a * d + b * c;
        }
        const int base = (i / p) * p * R + k;
        for (int q = 0; q < R; q++) {
          double sr = 0.0, si = 0.0;
          for (int r = 0; r < R; r++) {
            int ti = (int)(((long)r * q) % R) * m;
            double c = pl->wr[ti], d = sign < 0 ? pl->wi[ti] : -pl->wi[ti];
            sr += tr[r] * c - ti_[r] * d;
            si += tr[r] * d + ti_[r] * c;
          }
          dst[2 * (base + q * p)] = sr;
          dst[2 * (base + q * p) + 1] = si;
        }
        free(tr);
      }
    }
    double *t = src;
    src = dst;
    dst = t;
    p *= R;
  }
  if (src != x) memcpy(x, src, sizeof(double) * 2 * n);
}

/* Complex transform along one axis of an interleaved double array.  Pencil
 * (a, b) has element j at ((a * nb) + b) + j * stride, in complex units,
 * where the axis sits between the outer count na and the inner count nb.
 * Pencils are gathered AXIS_BLOCK at a time along b (adjacent in memory),
 * so every cache line read or written is used whole. */
#define AXIS_BLOCK 8
static void axis_pass(double *w, int n, int64_t na, int64_t nb, int64_t stride,
                      int64_t astride, int sign) {
  fplan pl;
  fplan_init(&pl, n);
  const int64_t nblk = (nb + AXIS_BLOCK - 1) / AXIS_BLOCK;
  const int64_t total = na * nblk;
#pragma omp parallel
  {
    double *x = (double *)malloc(sizeof(double) * 2 * n * (AXIS_BLOCK + 1));
    double *y = x + 2 * n * AXIS_BLOCK;
#pragma omp for schedule(static)
    for (int64_t pidx = 0; pidx < total; pidx++) {
      const int64_t a = pidx / nblk, b0 = (pidx % nblk) * AXIS_BLOCK;
      const int cnt = (int)(nb - b0 < AXIS_BLOCK ? nb - b0 : AXIS_BLOCK);
      double *base = w + 2 * (a * astride + b0);
      for (int j = 0; j < n; j++) {
        const double *src = base + 2 * (int64_t)j * stride;
        for (int c = 0; c < cnt; c++) {
          x[2 * (c * n + j)] = src[2 * c];
          x[2 * (c * n + j) + 1] = src[2 * c + 1];
        }
      }
      for (int c = 0; c < cnt; c++) fft_pencil(&pl, x + 2 * c * n, y, sign);
      for (int j = 0; j < n; j++) {
        double *dst = base + 2 * (int64_t)j * stride;
        for (int c = 0; c < cnt; c++) {
          dst[2 * c] = x[2 * (c * n + j)];
          dst[2 * c + 1] = x[2 * (c * n + j) + 1];
        }
      }
    }
    free(x);
  }
  fplan_free(&pl);
}

/* Forward unnormalised 3D real-to-half-complex DFT, dims (nz, ny, nx) with
 * x fastest, sign -1; output nz*ny*(nx/2+1) interleaved complex.
 * Restates Clarity_FFT_R2C_float (src/FFT.cxx:46-95). */
static int own_r2c(int nx, int ny, int nz, const float *in, float *out) {
  const int hx = nx / 2 + 1;
  const int64_t nrows = (int64_t)ny * nz;
  double *w = (double *)malloc(sizeof(double) * 2 * (size_t)nrows * hx);
  if (!w) return OR_OUT_OF_MEMORY;
  fplan pl;
  fplan_init(&pl, nx);
#pragma omp parallel
  {
    double *x = (double *)malloc(sizeof(double) * 4 * nx);
    double *y = x + 2 * nx;
#pragma omp for schedule(static)
    for (int64_t r = 0; r < nrows; r++) {
      const float *row = in + r * nx;
      for (int j = 0; j < nx; j++) {
        x[2 * j] = (double)row[j];
        x[2 * j + 1] = 0.0;
      }
      fft_pencil(&pl, x, y, -1);
      memcpy(w + 2 * r * hx, x, sizeof(double) * 2 * hx);
    }
    free(x);
  }
  fplan_free(&pl);
  /* y axis: pencils (z, kx), stride hx */
  axis_pass(w, ny, nz, hx, hx, (int64_t)ny * hx, -1);
  /* z axis: pencils (y, kx) -> treat as one outer block with nb = ny*hx */
  axis_pass(w, nz, 1, (int64_t)ny * hx, (int64_t)ny * hx, 0, -1);
  const int64_t nc = 2 * nrows * hx;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nc; i++) out[i] = (float)w[i];
  free(w);
  return OR_SUCCESS;
}

/* Inverse unnormalised 3D half-complex-to-real DFT, sign +1.  The imaginary
 * parts of kx = 0 and kx = nx/2 are ignored, as a multi-dimensional c2r
 * does: complex inverse along z and y, then a Hermitian real inverse along x.
 * Restates Clarity_FFT_C2R_float (src/FFT.cxx:98-146). */
static int own_c2r(int nx, int ny, int nz, const float *in, float *out) {
  const int hx = nx / 2 + 1;
  const int64_t nrows = (int64_t)ny * nz;
  const int64_t nc = 2 * nrows * hx;
  double *w = (double *)malloc(sizeof(double) * (size_t)nc);
  if (!w) return OR_OUT_OF_MEMORY;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nc; i++) w[i] = (double)in[i];
  axis_pass(w, nz, 1, (int64_t)ny * hx, (int64_t)ny * hx, 0, +1);
  axis_pass(w, ny, nz, hx, hx, (int64_t)ny * hx, +1);
  fplan pl;
  fplan_init(&pl, nx);
#pragma omp parallel
  {
    double *x = (double *)malloc(sizeof(double) * 4 * nx);
    double *y = x + 2 * nx;
#pragma omp for schedule(static)
    for (int64_t r = 0; r < nrows; r++) {
      const double *h = w + 2 * r * hx;
      x[0] = h[0];
      x[1] = 0.0;
      for (int k = 1; k < hx; k++) {
        x[2 * k] = h[2 * k];
        x[2 * k + 1] = h[2 * k + 1];
        if (nx - k != k) {
          x[2 * (nx - k)] = h[2 * k];
          x[2 * (nx - k) + 1] = -h[2 * k + 1];
        }
      }
      if (nx % 2 == 0) x[2 * (nx / 2) + 1] = 0.0;
      fft_pencil(&pl, x, y, +1);
      float *row = out + r * nx;
      for (int j = 0; j < nx; j++) row[j] = (float)x[2 * j];
    }
    free(x);
  }
  fplan_free(&pl);
  free(w);
  return OR_SUCCESS;
}

/* FFT entry points.  `in` of c2r may be destroyed by the FFTW backend (as
 * FFTW's multi-dimensional out-of-place c2r does), never by the own one. */
static int fft_r2c(int nx, int ny, int nz, const float *in, float *out) {
  if (g_fft_backend == ORACLE_FFT_FFTW_API && g_fw.loaded) {
    if (g_fw.nthreads) g_fw.nthreads(oracle_get_threads());
    /* dims reversed: (nz, ny, nx), src/FFT.cxx:76-81 */
    fw_plan p = g_fw.r2c(nz, ny, nx, (float *)in, out, FW_ESTIMATE);
    if (!p) return OR_FFT_FAILED;
    g_fw.exec(p);
    g_fw.destroy(p);
    return OR_SUCCESS;
  }
  return own_r2c(nx, ny, nz, in, out);
}

static int fft_c2r_destroy(int nx, int ny, int nz, float *in, float *out) {
  if (g_fft_backend == ORACLE_FFT_FFTW_API && g_fw.loaded) {
    if (g_fw.nthreads) g_fw.nthreads(oracle_get_threads());
    fw_plan p = g_fw.c2r(nz, ny, nx, in, out, FW_ESTIMATE);
    if (!p) return OR_FFT_FAILED;
    g_fw.exec(p);
    g_fw.destroy(p);
    return OR_SUCCESS;
  }
  return own_c2r(nx, ny, nz, in, out);
}

static int64_t complex_floats(int nx, int ny, int nz) {
  /* src/Memory.cxx:49: 2 * nz * ny * (nx/2 + 1) floats */
  return 2 * (int64_t)nz * ny * (nx / 2 + 1);
}

int oracle_fft_r2c(int nx, int ny, int nz, const float *in, float *out) {
  if (!in || !out || nx <= 0 || ny <= 0 || nz <= 0) return OR_INVALID_ARGUMENT;
  return fft_r2c(nx, ny, nz, in, out);
}

int oracle_fft_c2r(int nx, int ny, int nz, const float *in, float *out) {
  if (!in || !out || nx <= 0 || ny <= 0 || nz <= 0) return OR_INVALID_ARGUMENT;
  const int64_t nc = complex_floats(nx, ny, nz);
  float *tmp = (float *)malloc(sizeof(float) * (size_t)nc);
  if (!tmp) return OR_OUT_OF_MEMORY;
  memcpy(tmp, in, sizeof(float) * (size_t)nc);
  int rc = fft_c2r_destroy(nx, ny, nz, tmp, out);
  free(tmp);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Primitives                                                               */
/* ------------------------------------------------------------------------ */

/* Clarity_ReduceSum (src/ComputePrimitives.cxx:44-82).  GCC builds sum
 * sequentially in float32 (:61-71); the fp64 mode accumulates in double and
 * rounds once. */
float oracle_reduce_sum(const float *buf, int64_t n) {
  if (g_reduce_mode == ORACLE_REDUCE_SERIAL_F32) {
    float sum = 0.0f;
    for (int64_t i = 0; i < n; i++) sum += buf[i];
    return sum;
  }
  double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
  for (int64_t i = 0; i < n; i++) sum += (double)buf[i];
  return (float)sum;
}

/* Clarity_MultiplyArraysComponentWise (src/ComputePrimitives.cxx:85-118) */
void oracle_multiply(float *r, const float *a, const float *b, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) r[i] = a[i] * b[i];
}

/* Clarity_DivideArraysComponentWise (src/ComputePrimitives.cxx:121-158):
 * the guard compares fabs of the float divisor against the double 1e-5. */
void oracle_divide(float *r, const float *a, const float *b, float value,
                   int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    if ((double)fabsf(b[i]) < 1e-5)
      r[i] = value;
    else
      r[i] = a[i] / b[i];
  }
}

/* Clarity_ScaleArray (src/ComputePrimitives.cxx:161-194) */
void oracle_scale(float *r, const float *a, int64_t n, float scale) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) r[i] = scale * a[i];
}

/* Clarity_NormalizeArray (src/ComputePrimitives.cxx:197-209) */
void oracle_normalize(float *r, const float *a, int64_t n) {
  float sum = oracle_reduce_sum(a, n);
  oracle_scale(r, a, n, 1.0f / sum);
}

/* The 1/(nz*ny*nx) factor as the reference computes it: the expression
 * (float) nz*ny*nx evaluated left to right in float (src/Convolve.cxx:199,
 * src/WienerDeconvolve.cxx:54). */
static float inv_volume(int nx, int ny, int nz) {
  float v = (float)nz;
  v = v * (float)ny;
  v = v * (float)nx;
  return 1.0f / v;
}

/* Clarity_Modulate_KernelCPU (src/Convolve.cxx:195-205) with
 * ComplexMultiplyAndScale (include/internal/Complex.h:95-102). */
void oracle_modulate(int nx, int ny, int nz, const float *a, const float *b,
                     float *out) {
  const int64_t n = (int64_t)nz * ny * (nx / 2 + 1);
  const float scale = inv_volume(nx, ny, nz);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    float ar = a[2 * i], ai = a[2 * i + 1];
    float br = b[2 * i], bi = b[2 * i + 1];
    out[2 * i] = scale * (ar * br - ai * bi);
    out[2 * i + 1] = scale * (ar * bi + ai * br);
  }
}

/* Clarity_GetImageMax (src/JansenVanCittertDeconvolve.cxx:45-87) */
float oracle_image_max(const float *img, int64_t n) {
  float m = img[0];
  for (int64_t i = 0; i < n; i++)
    if (img[i] > m) m = img[i];
  return m;
}

/* JansenVanCittertDeconvolveKernelCPU
 * (src/JansenVanCittertDeconvolve.cxx:90-119) */
void oracle_jvc_update(int64_t n, const float *in, float a, float inv_a_sq,
                       const float *ik, const float *ok, float *out) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; j++) {
    float diff = ok[j] - a;
    float gamma = 1.0f - ((diff * diff) * inv_a_sq);
    float val = ik[j] + (gamma * (in[j] - ok[j]));
    if (val < 0.0f) val = 0.0f;
    out[j] = val;
  }
}

/* WienerDeconvolveKernelCPU (src/WienerDeconvolve.cxx:48-66) */
void oracle_wiener_kernel(int nx, int ny, int nz, const float *in_ft,
                          const float *psf_ft, float *out, float epsilon) {
  const int64_t n = (int64_t)nz * ny * (nx / 2 + 1);
  const float scale = inv_volume(nx, ny, nz);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    float hr = psf_ft[2 * i], hi = psf_ft[2 * i + 1];
    float cr = hr, ci = -hi;               /* ComplexConjugate */
    float mag = hr * hr + hi * hi;         /* ComplexMagnitudeSquared */
    float s = 1.0f / (mag + epsilon);
    cr = cr * s;                           /* ComplexMultiply(c, real) */
    ci = ci * s;
    float ar = in_ft[2 * i], ai = in_ft[2 * i + 1];
    out[2 * i] = scale * (cr * ar - ci * ai); /* ComplexMultiplyAndScale */
    out[2 * i + 1] = scale * (cr * ai + ci * ar);
  }
}

/* ------------------------------------------------------------------------ */
/* Pad / clip                                                               */
/* ------------------------------------------------------------------------ */

/* Clarity_ImagePadSpatialShift (src/ImagePadSpatialShift.cxx:32-69):
 * dst[d] = src[(d - shift) wrapped once then mod dstDim] when inside srcDim,
 * else fillValue.  The single += then % is kept as written. */
int oracle_pad_shift(float *dst, int dx, int dy, int dz, const float *src,
                     int sx, int sy, int sz, const int shift[3], float fill) {
  if (!dst || !src) return OR_INVALID_ARGUMENT;
  for (int dk = 0; dk < dz; dk++) {
    int sk = dk - shift[2];
    if (sk < 0) sk += dz;
    sk = sk % dz;
    int wk = sk >= 0 && sk < sz;
    for (int dj = 0; dj < dy; dj++) {
      int sj = dj - shift[1];
      if (sj < 0) sj += dy;
      sj = sj % dy;
      int wj = sj >= 0 && sj < sy;
      for (int di = 0; di < dx; di++) {
        int si = di - shift[0];
        if (si < 0) si += dx;
        si = si % dx;
        int wi = si >= 0 && si < sx;
        int64_t d = ((int64_t)dk * dy + dj) * dx + di;
        if (wi && wj && wk)
          dst[d] = src[((int64_t)sk * sy + sj) * sx + si];
        else
          dst[d] = fill;
      }
    }
  }
  return OR_SUCCESS;
}

/* Clarity_ImageClip (src/ImageClip.cxx:30-50): low-corner crop */
int oracle_clip(float *dst, int dx, int dy, int dz, const float *src, int sx,
                int sy, int sz) {
  (void)sz;
  if (!dst || !src) return OR_INVALID_ARGUMENT;
  for (int k = 0; k < dz; k++)
    for (int j = 0; j < dy; j++)
      for (int i = 0; i < dx; i++)
        dst[((int64_t)k * dy + j) * dx + i] = src[((int64_t)k * sy + j) * sx + i];
  return OR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Convolution and drivers                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
  int nx, ny, nz;
  int64_t v, c; /* real voxels, complex floats (2 per element) */
  float *in_pad, *k_pad;
} work_t;

static int dims_ok(int x, int y, int z) { return x > 0 && y > 0 && z > 0; }

/* work dims = image + kernel (src/MaximumLikelihoodDeconvolve.cxx:271-275),
 * observation in the low corner, kernel shifted by -k/2 with cyclic wrap
 * (:279-289). */
static int work_setup(work_t *w, const float *in, int ix, int iy, int iz,
                      const float *kernel, int kx, int ky, int kz) {
  memset(w, 0, sizeof *w);
  w->nx = ix + kx;
  w->ny = iy + ky;
  w->nz = iz + kz;
  w->v = (int64_t)w->nx * w->ny * w->nz;
  w->c = complex_floats(w->nx, w->ny, w->nz);
  w->in_pad = (float *)malloc(sizeof(float) * (size_t)w->v);
  w->k_pad = (float *)malloc(sizeof(float) * (size_t)w->v);
  if (!w->in_pad || !w->k_pad) return OR_OUT_OF_MEMORY;
  int zero[3] = {0, 0, 0};
  int kshift[3] = {-kx / 2, -ky / 2, -kz / 2};
  oracle_pad_shift(w->in_pad, w->nx, w->ny, w->nz, in, ix, iy, iz, zero, 0.0f);
  oracle_pad_shift(w->k_pad, w->nx, w->ny, w->nz, kernel, kx, ky, kz, kshift,
                   0.0f);
  return OR_SUCCESS;
}

static void work_free(work_t *w) {
  free(w->in_pad);
  free(w->k_pad);
}

/* Clarity_Convolve_OTF (src/Convolve.cxx:118-143): R2C, modulate, C2R,
 * with a fresh complex buffer per call as the reference allocates one. */
static int convolve_otf(const work_t *w, const float *in, const float *otf,
                        float *out) {
  float *ft = (float *)malloc(sizeof(float) * (size_t)w->c);
  if (!ft) return OR_OUT_OF_MEMORY;
  int rc = fft_r2c(w->nx, w->ny, w->nz, in, ft);
  if (rc == OR_SUCCESS) {
    oracle_modulate(w->nx, w->ny, w->nz, ft, otf, ft);
    rc = fft_c2r_destroy(w->nx, w->ny, w->nz, ft, out);
  }
  free(ft);
  return rc;
}

/* Clarity_Convolve (src/Convolve.cxx:41-115) via Clarity_ConvolveInternal
 * (:146-184). */
int oracle_convolve(const float *in, int ix, int iy, int iz,
                    const float *kernel, int kx, int ky, int kz, float *out) {
  if (!in || !kernel || !out) return OR_INVALID_ARGUMENT;
  if (!dims_ok(ix, iy, iz) || !dims_ok(kx, ky, kz)) return OR_INVALID_ARGUMENT;
  work_t w;
  int rc = work_setup(&w, in, ix, iy, iz, kernel, kx, ky, kz);
  float *in_ft = NULL, *k_ft = NULL, *res = NULL;
  if (rc == OR_SUCCESS) {
    in_ft = (float *)malloc(sizeof(float) * (size_t)w.c);
    k_ft = (float *)malloc(sizeof(float) * (size_t)w.c);
    res = (float *)malloc(sizeof(float) * (size_t)w.v);
    if (!in_ft || !k_ft || !res) rc = OR_OUT_OF_MEMORY;
  }
  if (rc == OR_SUCCESS) rc = fft_r2c(w.nx, w.ny, w.nz, w.in_pad, in_ft);
  if (rc == OR_SUCCESS) rc = fft_r2c(w.nx, w.ny, w.nz, w.k_pad, k_ft);
  if (rc == OR_SUCCESS) {
    oracle_modulate(w.nx, w.ny, w.nz, in_ft, k_ft, in_ft);
    rc = fft_c2r_destroy(w.nx, w.ny, w.nz, in_ft, res);
  }
  if (rc == OR_SUCCESS) oracle_clip(out, ix, iy, iz, res, w.nx, w.ny, w.nz);
  free(in_ft);
  free(k_ft);
  free(res);
  work_free(&w);
  return rc;
}

/* Clarity_WienerDeconvolve{,CPU} (src/WienerDeconvolve.cxx:69-122,211-290) */
int oracle_wiener(const float *in, int ix, int iy, int iz,
                  const float *kernel, int kx, int ky, int kz, float *out,
                  float epsilon) {
  if (!in || !kernel || !out) return OR_INVALID_ARGUMENT;
  if (!dims_ok(ix, iy, iz) || !dims_ok(kx, ky, kz)) return OR_INVALID_ARGUMENT;
  work_t w;
  int rc = work_setup(&w, in, ix, iy, iz, kernel, kx, ky, kz);
  float *in_ft = NULL, *k_ft = NULL, *res = NULL;
  if (rc == OR_SUCCESS) {
    in_ft = (float *)malloc(sizeof(float) * (size_t)w.c);
    k_ft = (float *)malloc(sizeof(float) * (size_t)w.c);
    res = (float *)malloc(sizeof(float) * (size_t)w.v);
    if (!in_ft || !k_ft || !res) rc = OR_OUT_OF_MEMORY;
  }
  if (rc == OR_SUCCESS) rc = fft_r2c(w.nx, w.ny, w.nz, w.in_pad, in_ft);
  if (rc == OR_SUCCESS) rc = fft_r2c(w.nx, w.ny, w.nz, w.k_pad, k_ft);
  if (rc == OR_SUCCESS) {
    oracle_wiener_kernel(w.nx, w.ny, w.nz, in_ft, k_ft, in_ft, epsilon);
    rc = fft_c2r_destroy(w.nx, w.ny, w.nz, in_ft, res);
  }
  if (rc == OR_SUCCESS) oracle_clip(out, ix, iy, iz, res, w.nx, w.ny, w.nz);
  free(in_ft);
  free(k_ft);
  free(res);
  work_free(&w);
  return rc;
}

/* Clarity_JansenVanCittertDeconvolve{,CPU}
 * (src/JansenVanCittertDeconvolve.cxx:122-192,311-389).  A = max/2 of the
 * unpadded input (:324, :128-129); i_0 = padded observation; iterations <= 1
 * still run one update (:167-173).  Negative counts are rejected here (the
 * reference would loop about 4e9 times). */
int oracle_jvc(const float *in, int ix, int iy, int iz, const float *kernel,
               int kx, int ky, int kz, float *out, int iterations) {
  if (!in || !kernel || !out) return OR_INVALID_ARGUMENT;
  if (!dims_ok(ix, iy, iz) || !dims_ok(kx, ky, kz)) return OR_INVALID_ARGUMENT;
  if (iterations < 0) return OR_INVALID_ARGUMENT;
  float mx = oracle_image_max(in, (int64_t)ix * iy * iz);
  work_t w;
  int rc = work_setup(&w, in, ix, iy, iz, kernel, kx, ky, kz);
  float a = 0.5f * mx;
  float inv_a_sq = 1.0f / (a * a);
  float *otf = NULL, *ik = NULL, *ok = NULL;
  if (rc == OR_SUCCESS) {
    otf = (float *)malloc(sizeof(float) * (size_t)w.c);
    ik = (float *)malloc(sizeof(float) * (size_t)w.v);
    ok = (float *)malloc(sizeof(float) * (size_t)w.v);
    if (!otf || !ik || !ok) rc = OR_OUT_OF_MEMORY;
  }
  if (rc == OR_SUCCESS) rc = fft_r2c(w.nx, w.ny, w.nz, w.k_pad, otf);
  if (rc == OR_SUCCESS) rc = convolve_otf(&w, w.in_pad, otf, ok);
  if (rc == OR_SUCCESS) {
    oracle_jvc_update(w.v, w.in_pad, a, inv_a_sq, w.in_pad, ok, ik);
    for (int k = 1; k < iterations && rc == OR_SUCCESS; k++) {
      rc = convolve_otf(&w, ik, otf, ok);
      if (rc == OR_SUCCESS)
        oracle_jvc_update(w.v, w.in_pad, a, inv_a_sq, ik, ok, ik);
    }
  }
  if (rc == OR_SUCCESS) oracle_clip(out, ix, iy, iz, ik, w.nx, w.ny, w.nz);
  free(otf);
  free(ik);
  free(ok);
  work_free(&w);
  return rc;
}

/* Clarity_MaximumLikelihoodUpdate (src/MaximumLikelihoodDeconvolve.cxx:48-81):
 * s1 = H (*) g; s1 = guard(obs / s1, 0); s2 = H (*) s1; g' = g * s2;
 * g' *= E / sum(g').  H (not conj(H)) in both convolutions. */
static int ml_update(const work_t *w, const float *obs, float energy,
                     const float *g, const float *otf, float *s1, float *s2,
                     float *gnew, float *scale_out) {
  int rc = convolve_otf(w, g, otf, s1);
  if (rc != OR_SUCCESS) return rc;
  oracle_divide(s1, obs, s1, 0.0f, w->v);
  rc = convolve_otf(w, s1, otf, s2);
  if (rc != OR_SUCCESS) return rc;
  oracle_multiply(gnew, g, s2, w->v);
  float new_energy = oracle_reduce_sum(gnew, w->v);
  float scale = energy / new_energy;
  if (scale_out) *scale_out = scale;
  oracle_scale(gnew, gnew, w->v, scale);
  return OR_SUCCESS;
}

/* Clarity_MaximumLikelihoodDeconvolve{,CPU}
 * (src/MaximumLikelihoodDeconvolve.cxx:84-153,255-332) with an optional
 * initial estimate (NULL: the padded observation, :137).  iterations == 0
 * copies the estimate and < 0 is rejected (the reference returns an
 * uninitialised buffer for 0 and loops ~4e9 times for < 0). */
int oracle_ml_ex(const float *in, int ix, int iy, int iz, const float *kernel,
                 int kx, int ky, int kz, const float *initial, float *out,
                 int iterations, float *scale_trace) {
  if (!in || !kernel || !out) return OR_INVALID_ARGUMENT;
  if (!dims_ok(ix, iy, iz) || !dims_ok(kx, ky, kz)) return OR_INVALID_ARGUMENT;
  if (iterations < 0) return OR_INVALID_ARGUMENT;
  work_t w;
  int rc = work_setup(&w, in, ix, iy, iz, kernel, kx, ky, kz);
  const double td0 = now_seconds();
  float *otf = NULL, *g = NULL, *s1 = NULL, *s2 = NULL;
  if (rc == OR_SUCCESS) {
    otf = (float *)malloc(sizeof(float) * (size_t)w.c);
    g = (float *)malloc(sizeof(float) * (size_t)w.v);
    s1 = (float *)malloc(sizeof(float) * (size_t)w.v);
    s2 = (float *)malloc(sizeof(float) * (size_t)w.v);
    if (!otf || !g || !s1 || !s2) rc = OR_OUT_OF_MEMORY;
  }
  if (rc == OR_SUCCESS) {
    if (initial) {
      int zero[3] = {0, 0, 0};
      oracle_pad_shift(g, w.nx, w.ny, w.nz, initial, ix, iy, iz, zero, 0.0f);
    } else {
      memcpy(g, w.in_pad, sizeof(float) * (size_t)w.v);
    }
  }
  if (rc == OR_SUCCESS && iterations > 0) rc = fft_r2c(w.nx, w.ny, w.nz, w.k_pad, otf);
  g_loop_seconds = 0.0;
  if (rc == OR_SUCCESS && iterations > 0) {
    float energy = oracle_reduce_sum(w.in_pad, w.v);
    double t0 = now_seconds();
    for (int k = 0; k < iterations && rc == OR_SUCCESS; k++)
      rc = ml_update(&w, w.in_pad, energy, g, otf, s1, s2, g,
                     scale_trace ? scale_trace + k : NULL);
    g_loop_seconds = now_seconds() - t0;
  }
  g_deconv_seconds = now_seconds() - td0;
  if (rc == OR_SUCCESS) oracle_clip(out, ix, iy, iz, g, w.nx, w.ny, w.nz);
  free(otf);
  free(g);
  free(s1);
  free(s2);
  work_free(&w);
  return rc;
}

int oracle_ml(const float *in, int ix, int iy, int iz, const float *kernel,
              int kx, int ky, int kz, float *out, int iterations) {
  return oracle_ml_ex(in, ix, iy, iz, kernel, kx, ky, kz, NULL, out,
                      iterations, NULL);
}

/* Clarity_BlindMaximumLikelihoodUpdate / ...DeconvolveCPU as written
 * (src/BlindMaximumLikelihoodDeconvolve.cxx:50-99,102-148,249-317):
 * normalise h in place (reduce + scale by 1/sum), o' = o * (H (*) guard(obs /
 * (H (*) o))), no energy rescale, h otherwise unchanged (the PSF update at
 * :96 is commented out; the third modulate/C2R at :89-93 has no effect and is
 * skipped here). */
int oracle_blind_ml(const float *in, int ix, int iy, int iz,
                    const float *kernel, int kx, int ky, int kz, float *out,
                    int iterations) {
  if (!in || !kernel || !out) return OR_INVALID_ARGUMENT;
  if (!dims_ok(ix, iy, iz) || !dims_ok(kx, ky, kz)) return OR_INVALID_ARGUMENT;
  if (iterations < 0) return OR_INVALID_ARGUMENT;
  work_t w;
  int rc = work_setup(&w, in, ix, iy, iz, kernel, kx, ky, kz);
  float *hft = NULL, *o = NULL, *tmp = NULL;
  if (rc == OR_SUCCESS) {
    hft = (float *)malloc(sizeof(float) * (size_t)w.c);
    o = (float *)malloc(sizeof(float) * (size_t)w.v);
    tmp = (float *)malloc(sizeof(float) * (size_t)w.v);
    if (!hft || !o || !tmp) rc = OR_OUT_OF_MEMORY;
  }
  if (rc == OR_SUCCESS) memcpy(o, w.in_pad, sizeof(float) * (size_t)w.v);
  for (int k = 0; k < iterations && rc == OR_SUCCESS; k++) {
    oracle_normalize(w.k_pad, w.k_pad, w.v);
    rc = fft_r2c(w.nx, w.ny, w.nz, w.k_pad, hft);
    if (rc == OR_SUCCESS) rc = convolve_otf(&w, o, hft, tmp);
    if (rc != OR_SUCCESS) break;
    oracle_divide(tmp, w.in_pad, tmp, 0.0f, w.v);
    rc = convolve_otf(&w, tmp, hft, tmp);
    if (rc == OR_SUCCESS) oracle_multiply(o, o, tmp, w.v);
  }
  if (rc == OR_SUCCESS) oracle_clip(out, ix, iy, iz, o, w.nx, w.ny, w.nz);
  free(hft);
  free(o);
  free(tmp);
  work_free(&w);
  return rc;
}

/* Alternating blind RL (no counterpart in the reference; it completes the
 * PSF step the reference leaves commented out at
 * src/BlindMaximumLikelihoodDeconvolve.cxx:89-96).  Per outer iteration:
 * `inner` object RL updates with energy rescale (as ml_update, H in both
 * convolutions, guard 0), then `inner` PSF updates
 *   h'(z) = h(z) * sum_y r(y) o(y - z),  r = guard(obs / (h (*) o)),
 * i.e. the correction is the CORRELATION of r with o (spectrum R conj(O)),
 * each followed by renormalising h to sum 1 in float64.  This is the
 * Richardson-Lucy step for h with o fixed (the multiplicative EM update of
 * the Poisson likelihood in h); the commented-out reference line modulates
 * with O itself (a convolution), which is that step only for an object
 * symmetric about the origin: with the true object and PSF it moves the
 * PSF by O(1), where the correlation leaves it fixed.  h starts as the
 * padded, shifted kernel normalised to sum 1 in float64; o as `initial`
 * (padded, image dims) or, NULL, the padded observation.  psf_out
 * (optional) receives h un-shifted and clipped to the kernel dims. */
static void normalize_fp64(float *h, int64_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; i++) s += (double)h[i];
  float inv = (float)(1.0 / s);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) h[i] = h[i] * inv;
}

/* the correction's correlation: C2R(R2C(in) * conj(ft) * scale), the
 * product in ComplexMultiplyAndScale's order with conj(b) = (c, -d) */
static int correlate_otf(const work_t *w, const float *in, const float *ft,
                         float *out) {
  float *spec = (float *)malloc(sizeof(float) * (size_t)w->c);
  if (!spec) return OR_OUT_OF_MEMORY;
  int rc = fft_r2c(w->nx, w->ny, w->nz, in, spec);
  if (rc == OR_SUCCESS) {
    const int64_t n = w->c / 2;
    const float scale = inv_volume(w->nx, w->ny, w->nz);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      float ar = spec[2 * i], ai = spec[2 * i + 1];
      float br = ft[2 * i], bi = -ft[2 * i + 1];
      spec[2 * i] = scale * (ar * br - ai * bi);
      spec[2 * i + 1] = scale * (ar * bi + ai * br);
    }
    rc = fft_c2r_destroy(w->nx, w->ny, w->nz, spec, out);
  }
  free(spec);
  return rc;
}

int oracle_blind_rl(const float *in, int ix, int iy, int iz,
                    const float *kernel, int kx, int ky, int kz,
                    const float *initial, float *out, float *psf_out,
                    int outer, int inner) {
  if (!in || !kernel || !out) return OR_INVALID_ARGUMENT;
  if (!dims_ok(ix, iy, iz) || !dims_ok(kx, ky, kz)) return OR_INVALID_ARGUMENT;
  if (outer < 0 || inner < 0) return OR_INVALID_ARGUMENT;
  work_t w;
  int rc = work_setup(&w, in, ix, iy, iz, kernel, kx, ky, kz);
  float *hft = NULL, *oft = NULL, *o = NULL, *s1 = NULL, *s2 = NULL;
  if (rc == OR_SUCCESS) {
    hft = (float *)malloc(sizeof(float) * (size_t)w.c);
    oft = (float *)malloc(sizeof(float) * (size_t)w.c);
    o = (float *)malloc(sizeof(float) * (size_t)w.v);
    s1 = (float *)malloc(sizeof(float) * (size_t)w.v);
    s2 = (float *)malloc(sizeof(float) * (size_t)w.v);
    if (!hft || !oft || !o || !s1 || !s2) rc = OR_OUT_OF_MEMORY;
  }
  float *h = w.k_pad;
  if (rc == OR_SUCCESS) {
    if (initial) {
      const int zero[3] = {0, 0, 0};
      rc = oracle_pad_shift(o, w.nx, w.ny, w.nz, initial, ix, iy, iz, zero, 0.0f);
    } else {
      memcpy(o, w.in_pad, sizeof(float) * (size_t)w.v);
    }
    normalize_fp64(h, w.v);
  }
  float energy = rc == OR_SUCCESS ? oracle_reduce_sum(w.in_pad, w.v) : 0.0f;
  for (int it = 0; it < outer && rc == OR_SUCCESS; it++) {
    rc = fft_r2c(w.nx, w.ny, w.nz, h, hft);
    for (int k = 0; k < inner && rc == OR_SUCCESS; k++)
      rc = ml_update(&w, w.in_pad, energy, o, hft, s1, s2, o, NULL);
    if (rc == OR_SUCCESS) rc = fft_r2c(w.nx, w.ny, w.nz, o, oft);
    for (int k = 0; k < inner && rc == OR_SUCCESS; k++) {
      rc = convolve_otf(&w, h, oft, s1);
      if (rc != OR_SUCCESS) break;
      oracle_divide(s1, w.in_pad, s1, 0.0f, w.v);
      rc = correlate_otf(&w, s1, oft, s2);
      if (rc != OR_SUCCESS) break;
      oracle_multiply(h, h, s2, w.v);
      normalize_fp64(h, w.v);
    }
  }
  if (rc == OR_SUCCESS) oracle_clip(out, ix, iy, iz, o, w.nx, w.ny, w.nz);
  if (rc == OR_SUCCESS && psf_out) {
    /* inverse of the -k/2 cyclic shift, then the kernel-sized low corner */
    for (int k = 0; k < kz; k++)
      for (int j = 0; j < ky; j++)
        for (int i = 0; i < kx; i++) {
          int sk = ((k - kz / 2) % w.nz + w.nz) % w.nz;
          int sj = ((j - ky / 2) % w.ny + w.ny) % w.ny;
          int si = ((i - kx / 2) % w.nx + w.nx) % w.nx;
          psf_out[((int64_t)k * ky + j) * kx + i] =
              h[((int64_t)sk * w.ny + sj) * w.nx + si];
        }
  }
  free(hft);
  free(oft);
  free(o);
  free(s1);
  free(s2);
  work_free(&w);
  return rc;
}

/* IDIVFunctional (src/IDivergenceDeconvolve.cxx:33-47) with alpha = 0,
 * accumulated in float64, and 0 * ln 0 taken as 0 (the reference gives NaN
 * for g = 0). */
double oracle_idivergence(const float *g, const float *ghat, int64_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double a = g[i], b = ghat[i];
    double t = (a > 0.0) ? a * log(a / b) : 0.0;
    s += t + b - a;
  }
  return s;
}
