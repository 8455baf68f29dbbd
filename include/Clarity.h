/*
 * Clarity.h -- drop-in C API of the MI355X Clarity deconvolution library.
 *
 * Same names, argument meaning, enum values and struct layout as the
 * reference's public header (Clarity.h.in), so a program written against the
 * reference compiles and links against libclarity.so unchanged.  Each
 * declaration cites the reference declaration it replaces.
 *
 * Semantics shared by every image function:
 *   - volumes are float32, x fastest: index (z*ny + y)*nx + x
 *     (reference src/ImageClip.cxx:41);
 *   - all image pointers are caller-owned HOST buffers; inputs are not
 *     modified; outputs have imageDim size;
 *   - the FFT work domain is imageDim + kernelDim per axis, the observation
 *     sits in the low corner and the kernel is cyclically shifted by -k/2
 *     (reference src/MaximumLikelihoodDeconvolve.cxx:271-289);
 *   - calls are synchronous: results are in host memory on return.
 * All compute runs on the GPU (HIP, gfx950); there is no CPU code path.
 */
#ifndef __CLARITY_LIB_H_
#define __CLARITY_LIB_H_

/* Clarity.h.in:41 -- version 1.1.0 */
#define CLARITY_VERSION 110

#if defined(__GNUC__) || defined(__clang__)
#define CLARITY_EXPORT __attribute__((visibility("default")))
#else
#define CLARITY_EXPORT
#endif

#ifdef __cplusplus
#define C_FUNC_DEF extern "C" CLARITY_EXPORT
#else
#define C_FUNC_DEF CLARITY_EXPORT
#endif

#ifndef NULL
#define NULL 0L
#endif

/* Clarity.h.in:90-101 -- values kept: CLARITY_SUCCESS is 5, not 0. */
typedef enum {
  CLARITY_FFT_FAILED,           /* FFT plan or execution failed */
  CLARITY_OUT_OF_MEMORY,        /* host allocation failed */
  CLARITY_DEVICE_OUT_OF_MEMORY, /* device (HBM) allocation failed */
  CLARITY_INVALID_OPERATION,    /* no usable GPU, or a GPU call failed */
  CLARITY_INVALID_ARGUMENT,     /* NULL pointer, dims <= 0, bad count */
  CLARITY_SUCCESS
} ClarityResult_t;

/* Clarity.h.in:110-114 -- three ints by value, x then y then z. */
typedef struct _Clarity_Dim3_t {
  int x;
  int y;
  int z;
} Clarity_Dim3;

/* Clarity.h.in:123-125 */
C_FUNC_DEF Clarity_Dim3 Clarity_Dim3FromArray(int dimArray[3]);

/* Clarity.h.in:137-139 (reference src/Clarity.cxx:46-79).  Reference
 * counted; the first call selects HIP device 0 (or $CLARITY_DEVICE) and
 * creates the library stream.  Returns CLARITY_INVALID_OPERATION when no HIP
 * device is usable.  As with the reference, calling it is optional: an entry
 * point used before any Register registers implicitly (one client, held until
 * an UnRegister or process exit). */
C_FUNC_DEF ClarityResult_t Clarity_Register();

/* Clarity.h.in:147-149 (reference src/Clarity.cxx:82-91).  The last matching
 * call releases cached device memory, FFT plans and the stream. */
C_FUNC_DEF ClarityResult_t Clarity_UnRegister();

/* The name the reference used before Clarity_Register (reference
 * include/ChangeLog.txt:82); aliases of Register / UnRegister. */
C_FUNC_DEF ClarityResult_t Clarity_Init();
C_FUNC_DEF ClarityResult_t Clarity_Shutdown();

/* Clarity.h.in:160-162 (reference src/Clarity.cxx:93-103).  Host thread
 * count; recorded only, since no host thread does compute here. */
C_FUNC_DEF ClarityResult_t Clarity_SetNumberOfThreads(unsigned n);

/* Clarity.h.in:181-185 (reference src/ImagePadSpatialShift.cxx:32-69):
 * dst[d] = src[(d - shift) cyclically in dstDim] when that lands inside
 * srcDim, else fillValue.  Runs on the GPU. */
C_FUNC_DEF ClarityResult_t Clarity_ImagePadSpatialShift(
    float *dst, Clarity_Dim3 dstDim, float *src, Clarity_Dim3 srcDim,
    int shift[3], float fillValue);

/* Clarity.h.in:200-202 (reference src/ImageClip.cxx:30-50): copies the
 * [0,dstDim) low corner of src.  Runs on the GPU. */
C_FUNC_DEF ClarityResult_t Clarity_ImageClip(float *dst, Clarity_Dim3 dstDim,
                                             float *src, Clarity_Dim3 srcDim);

/* Clarity.h.in:210-212 (reference src/Clarity.cxx:116-150): static string. */
C_FUNC_DEF char *Clarity_ResultString(ClarityResult_t result);

/* Clarity.h.in:236-240 (reference src/WienerDeconvolve.cxx:211-290):
 * out = C2R( conj(H)/(|H|^2 + epsilon) * FFT(in) ) / N. */
C_FUNC_DEF ClarityResult_t Clarity_WienerDeconvolve(float *inImage,
                                                    Clarity_Dim3 imageDim,
                                                    float *kernelImage,
                                                    Clarity_Dim3 kernelDim,
                                                    float *outImage,
                                                    float epsilon);

/* Clarity.h.in:257-261 (reference src/JansenVanCittertDeconvolve.cxx:
 * 311-389).  iterations 0 and 1 both run one update, as the reference does;
 * iterations < 0 returns CLARITY_INVALID_ARGUMENT. */
C_FUNC_DEF ClarityResult_t Clarity_JansenVanCittertDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage, int iterations);

/* Clarity.h.in:285-291.  The reference implementation is a stub that writes
 * zeros (reference src/SmoothedJansenVanCittertDeconvolve.cxx:28-39); the
 * same stub is kept for ABI completeness. */
C_FUNC_DEF ClarityResult_t Clarity_SmoothedJansenVanCittertDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage, unsigned iterations,
    unsigned smoothInterval, float *smoothSigma[3]);

/* Clarity.h.in:314-319.  Reference stub (src/IDivergenceDeconvolve.cxx:
 * 59-72) writing zeros; kept as such.  The I-divergence functional itself is
 * available as ClarityEx_IDivergence. */
C_FUNC_DEF ClarityResult_t Clarity_IDivergenceDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage);

/* Clarity.h.in:337-341 (reference src/MaximumLikelihoodDeconvolve.cxx:
 * 255-332).  Richardson-Lucy: g0 = observation; per iteration
 * g' = g * (H (*) guard(obs / (H (*) g))), then g' *= E_obs / sum(g'), with
 * H (not conj H) in both convolutions and guard |s| < 1e-5 -> 0.  Sums are
 * float64 on the device.  iterations == 0 returns the observation;
 * iterations < 0 returns CLARITY_INVALID_ARGUMENT. */
C_FUNC_DEF ClarityResult_t Clarity_MaximumLikelihoodDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage, int iterations);

/* Clarity.h.in:358-363 (reference src/BlindMaximumLikelihoodDeconvolve.cxx:
 * 249-317, as written: the kernel is renormalised every iteration, the image
 * updated as RL without energy rescale, the kernel never updated).  For a
 * blind method that does update the kernel see ClarityEx_BlindRL. */
C_FUNC_DEF ClarityResult_t Clarity_BlindMaximumLikelihoodDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage, int iterations);

/* Clarity.h.in:379-383 (reference src/Convolve.cxx:41-115). */
C_FUNC_DEF ClarityResult_t Clarity_Convolve(float *inImage,
                                            Clarity_Dim3 imageDim,
                                            float *kernel,
                                            Clarity_Dim3 kernelDim,
                                            float *outImage);

/* Internal-but-extern-"C" FFT entry points of the reference
 * (include/internal/FFT.h:40-58, src/FFT.cxx:46-146).  Unnormalised 3D
 * real <-> half-complex DFT with FFTW's dim order (nz, ny, nx): in/out are
 * HOST buffers of nz*ny*nx floats and nz*ny*(nx/2+1) interleaved complex
 * floats.  Forward sign -1, inverse +1; C2R ignores the imaginary parts of
 * the kx = 0 and kx = nx/2 planes and, unlike FFTW, leaves its input intact. */
C_FUNC_DEF ClarityResult_t Clarity_FFT_R2C_float(int nx, int ny, int nz,
                                                 float *in, float *out);
C_FUNC_DEF ClarityResult_t Clarity_FFT_C2R_float(int nx, int ny, int nz,
                                                 float *in, float *out);

#endif /* __CLARITY_LIB_H_ */
