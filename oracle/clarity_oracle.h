/*
 * clarity_oracle.h -- CPU restatement of Clarity's FFT deconvolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against (and the CPU baseline that bench.py times).  Nothing in the product
 * (clarity_amd/, include/) links, loads or calls it.
 *
 * Every function restates one function of the reference (paths relative to
 * the reference tree) with the same argument meaning, element order and
 * float32 operation order.  Two switches make the remaining freedom explicit:
 *
 *   reduce mode  ORACLE_REDUCE_SERIAL_F32  sequential float32 sum, i = 0..n-1,
 *                                          as the reference's GCC build does
 *                                          (src/ComputePrimitives.cxx:61-71)
 *                ORACLE_REDUCE_FP64        double accumulation, rounded to
 *                                          float once (the HIP library's rule)
 *   fft backend  ORACLE_FFT_OWN            own mixed-radix FFT in float64,
 *                                          rounded to float32 at the end
 *                ORACLE_FFT_FFTW_API       the FFTW3 single-precision API
 *                                          (fftwf_plan_dft_*_3d + execute with
 *                                          FFTW_ESTIMATE and a plan per call,
 *                                          as src/FFT.cxx:80-86,131-137),
 *                                          resolved with dlopen at run time
 *                                          from libfftw3f or MKL's FFTW3
 *                                          interface.  Used for timing.
 *
 * Result codes use the reference's ClarityResult_t values
 * (Clarity.h.in:90-101): SUCCESS is 5.
 */
#ifndef CLARITY_ORACLE_H
#define CLARITY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OR_FFT_FAILED = 0,
  OR_OUT_OF_MEMORY = 1,
  OR_DEVICE_OUT_OF_MEMORY = 2,
  OR_INVALID_OPERATION = 3,
  OR_INVALID_ARGUMENT = 4,
  OR_SUCCESS = 5
};

enum { ORACLE_REDUCE_SERIAL_F32 = 0, ORACLE_REDUCE_FP64 = 1 };
enum { ORACLE_FFT_OWN = 0, ORACLE_FFT_FFTW_API = 1 };

/* configuration */
void oracle_set_reduce_mode(int mode);
int  oracle_get_reduce_mode(void);
int  oracle_set_fft_backend(int backend);     /* returns OR_SUCCESS or OR_FFT_FAILED */
int  oracle_get_fft_backend(void);
const char *oracle_fft_backend_name(void);
void oracle_set_threads(int n);
int  oracle_get_threads(void);

/* FFT (src/FFT.cxx) */
int oracle_fft_r2c(int nx, int ny, int nz, const float *in, float *out);
int oracle_fft_c2r(int nx, int ny, int nz, const float *in, float *out);

/* primitives (src/ComputePrimitives.cxx, src/Convolve.cxx) */
float oracle_reduce_sum(const float *buf, int64_t n);
void  oracle_multiply(float *r, const float *a, const float *b, int64_t n);
void  oracle_divide(float *r, const float *a, const float *b, float value,
                    int64_t n);
void  oracle_scale(float *r, const float *a, int64_t n, float scale);
void  oracle_normalize(float *r, const float *a, int64_t n);
void  oracle_modulate(int nx, int ny, int nz, const float *a, const float *b,
                      float *out);
float oracle_image_max(const float *img, int64_t n);
void  oracle_jvc_update(int64_t n, const float *in, float a, float inv_a_sq,
                        const float *ik, const float *ok, float *out);
void  oracle_wiener_kernel(int nx, int ny, int nz, const float *in_ft,
                           const float *psf_ft, float *out, float epsilon);

/* pad / clip (src/ImagePadSpatialShift.cxx, src/ImageClip.cxx) */
int oracle_pad_shift(float *dst, int dx, int dy, int dz, const float *src,
                     int sx, int sy, int sz, const int shift[3], float fill);
int oracle_clip(float *dst, int dx, int dy, int dz, const float *src, int sx,
                int sy, int sz);

/* public drivers (image/kernel dims given as x,y,z) */
int oracle_convolve(const float *in, int ix, int iy, int iz,
                    const float *kernel, int kx, int ky, int kz, float *out);
int oracle_wiener(const float *in, int ix, int iy, int iz,
                  const float *kernel, int kx, int ky, int kz, float *out,
                  float epsilon);
int oracle_jvc(const float *in, int ix, int iy, int iz, const float *kernel,
               int kx, int ky, int kz, float *out, int iterations);
int oracle_ml(const float *in, int ix, int iy, int iz, const float *kernel,
              int kx, int ky, int kz, float *out, int iterations);
int oracle_blind_ml(const float *in, int ix, int iy, int iz,
                    const float *kernel, int kx, int ky, int kz, float *out,
                    int iterations);

/* extensions: RL from a caller-given estimate (NULL = the observation) with
 * an optional per-iteration trace of the rescale factor; alternating blind
 * RL; I-divergence. */
int oracle_ml_ex(const float *in, int ix, int iy, int iz, const float *kernel,
                 int kx, int ky, int kz, const float *initial, float *out,
                 int iterations, float *scale_trace);
int oracle_blind_rl(const float *in, int ix, int iy, int iz,
                    const float *kernel, int kx, int ky, int kz,
                    const float *initial, float *out, float *psf_out,
                    int outer, int inner);
double oracle_idivergence(const float *g, const float *ghat, int64_t n);

/* Wall seconds of the iteration loop of the last oracle_ml / oracle_ml_ex
 * call (the reference's TIME_DECONVOLVE scope minus the OTF and energy). */
double oracle_last_loop_seconds(void);
/* Wall seconds of the last RL call's deconvolution scope (after padding,
 * before clipping). */
double oracle_last_deconvolve_seconds(void);

#ifdef __cplusplus
}
#endif
#endif
