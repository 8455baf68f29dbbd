/*
 * ClarityEx.h -- extensions of the MI355X Clarity library.
 *
 * None of these has a counterpart in the reference's public header; each
 * comment names the reference behaviour it builds on.  Host-pointer
 * functions are synchronous like the Clarity_* API.  The plan API keeps
 * every buffer resident in HBM between calls, for callers that deconvolve
 * many times on one shape (benchmarks, pipelines, warm starts).
 */
#ifndef __CLARITY_EX_H_
#define __CLARITY_EX_H_

#include <stdint.h>

#include "Clarity.h"

/* Richardson-Lucy from a caller-given initial estimate (imageDim floats,
 * NULL = the observation, which is the reference's start,
 * src/MaximumLikelihoodDeconvolve.cxx:137).  Everything else as
 * Clarity_MaximumLikelihoodDeconvolve.  Used by warm starts and to resume. */
C_FUNC_DEF ClarityResult_t ClarityEx_MaximumLikelihoodDeconvolveFrom(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *initialImage, float *outImage,
    int iterations);

/* Alternating blind Richardson-Lucy.  Completes the kernel step that the
 * reference leaves commented out (src/BlindMaximumLikelihoodDeconvolve.cxx:
 * 89-96).  Per outer iteration: `inner` image RL updates exactly as
 * Clarity_MaximumLikelihoodDeconvolve (H in both convolutions, guard -> 0,
 * energy rescale), then `inner` kernel updates
 *   h'(z) = h(z) * sum_y r(y) O(y - z),   r = guard(obs / (h (*) O)),
 * the Richardson-Lucy step for h with the image O fixed (the correction is
 * the correlation of r with O, spectrum R conj(O); the commented-out
 * reference line modulates with O itself, a convolution, which moves even
 * the true kernel), each followed by renormalising h to sum 1 in float64.
 * kernelOut (kernelDim floats, may be NULL) receives the final kernel. */
C_FUNC_DEF ClarityResult_t ClarityEx_BlindRL(float *inImage,
                                             Clarity_Dim3 imageDim,
                                             float *kernelImage,
                                             Clarity_Dim3 kernelDim,
                                             float *outImage,
                                             float *kernelOut, int outer,
                                             int inner);

/* ClarityEx_BlindRL from a caller-given first image estimate (imageDim
 * floats; NULL = the observation, ClarityEx_BlindRL's start). */
C_FUNC_DEF ClarityResult_t ClarityEx_BlindRLFrom(float *inImage,
                                                 Clarity_Dim3 imageDim,
                                                 float *kernelImage,
                                                 Clarity_Dim3 kernelDim,
                                                 float *initialImage,
                                                 float *outImage,
                                                 float *kernelOut, int outer,
                                                 int inner);

/* I-divergence sum_i g ln(g/ghat) + ghat - g over n host floats, float64 on
 * the device, 0 ln 0 := 0 (reference functional src/IDivergenceDeconvolve.cxx:
 * 33-47 with alpha = 0). */
C_FUNC_DEF ClarityResult_t ClarityEx_IDivergence(const float *g,
                                                 const float *ghat, int64_t n,
                                                 double *result);

/* Richardson-Lucy on `count` independent volumes of (possibly) different
 * sizes, each with its own kernel, `iterations` updates each; every volume
 * gets exactly the result of Clarity_MaximumLikelihoodDeconvolve on it.  The
 * volumes run concurrently on several HIP streams (small volumes alone do
 * not fill an MI355X).  Arrays of host pointers and dims, `count` long. */
C_FUNC_DEF ClarityResult_t ClarityEx_MaximumLikelihoodDeconvolveBatch(
    int count, float *const *inImages, const Clarity_Dim3 *imageDims,
    float *const *kernelImages, const Clarity_Dim3 *kernelDims,
    float *const *outImages, int iterations);

/* ---- resident Richardson-Lucy plan ------------------------------------ */

typedef struct ClarityEx_RLPlan_t *ClarityEx_RLPlan;

/* Creates a plan for one (imageDim, kernelDim) pair: allocates the HBM
 * buffers and FFT twiddles.  No data yet.  The plan holds a registration
 * (Clarity_Register) until ClarityEx_RLPlanDestroy, so a caller's last
 * Clarity_UnRegister does not free what the plan uses. */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanCreate(ClarityEx_RLPlan *plan,
                                                  Clarity_Dim3 imageDim,
                                                  Clarity_Dim3 kernelDim);

/* Loads the observation (imageDim floats) and kernel (kernelDim floats) and
 * computes the OTF.  onDevice != 0: both pointers are device pointers. */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanSetInputs(ClarityEx_RLPlan plan,
                                                     const float *inImage,
                                                     const float *kernelImage,
                                                     int onDevice);

/* Sets the estimate to `initial` (imageDim floats; NULL = observation). */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanReset(ClarityEx_RLPlan plan,
                                                 const float *initial,
                                                 int onDevice);

/* Runs `iterations` RL updates on the resident estimate, asynchronously on
 * the library stream (no host synchronisation inside). */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanIterate(ClarityEx_RLPlan plan,
                                                   int iterations);

/* Copies the current (energy-scaled) estimate out; waits for the stream. */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanGetEstimate(ClarityEx_RLPlan plan,
                                                       float *out,
                                                       int onDevice);

/* Last rescale factor E_obs / sum(g') (float), read back from HBM. */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanGetScale(ClarityEx_RLPlan plan,
                                                    float *scale);

C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanDestroy(ClarityEx_RLPlan plan);

/* Describes the plan's kernels: writes up to `cap` entries, returns count.
 * names[i] is a static string, bytes[i] the algorithmic HBM bytes one launch
 * of that kernel moves, launches[i] its launches per RL iteration of a long
 * run (kernels that run once per ClarityEx_RLPlanIterate call report 0). */
C_FUNC_DEF int ClarityEx_RLPlanDescribe(ClarityEx_RLPlan plan,
                                        const char **names, double *bytes,
                                        int *launches, int cap);

/* Per-kernel timing with HIP events recorded on the library stream around
 * each launch (0 = off, the default).  ClarityEx_RLPlanTimings returns for
 * each kernel of ClarityEx_RLPlanDescribe the launch count and summed
 * milliseconds since the last call with reset != 0. */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanSetTiming(ClarityEx_RLPlan plan,
                                                     int enable);
C_FUNC_DEF int ClarityEx_RLPlanTimings(ClarityEx_RLPlan plan, int *count,
                                       double *ms, int cap, int reset);

/* Chooses the fused engine's launch shapes (tile sizes, elements per thread,
 * pencils per workgroup, chaining, specialisations) for this plan's shape by
 * timing candidates on its inputs, `iterations` iterations each (coordinate
 * descent from the built-in defaults).  The estimate is saved and restored,
 * so results are unchanged; later plans of the same shape in the process use
 * the choice (unless CLARITY_FAST_TUNE is set).  Writes the chosen shapes, in
 * CLARITY_FAST_TUNE syntax, to `description` (NUL-terminated, at most `cap`
 * bytes; may be NULL).  No counterpart in the reference.  Requires inputs. */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanAutotune(ClarityEx_RLPlan plan, int iterations,
                                                    char *description, int cap);

/* Name of the engine running the plan: "fast" (fused power-of-two
 * pipeline) or "generic". */
C_FUNC_DEF const char *ClarityEx_RLPlanEngine(ClarityEx_RLPlan plan);

/* The FFT work domain the plan's engine transforms over.  The generic
 * engine uses the reference's imageDim + kernelDim per axis
 * (src/MaximumLikelihoodDeconvolve.cxx:271-289); the fused engine the
 * smallest of its sizes >= max(kernel, image + kernel/2), over which the RL
 * iterates in the image are the same (nothing outside the image is kept,
 * and the kernel cannot wrap into the offsets the image uses), with nz = 2
 * image z when image and kernel share their depth.  CLARITY_REFERENCE_DOMAIN
 * makes the fused engine keep imageDim + kernelDim where it can. */
C_FUNC_DEF ClarityResult_t ClarityEx_RLPlanWorkDims(ClarityEx_RLPlan plan,
                                                    Clarity_Dim3 *workDim);

/* ---- learned warm start ---------------------------------------------- */

/* A 3-level 3D U-Net (channels 16/32/64, 3x3x3 convolutions, ReLU, skip
 * connections; bfloat16 weights on MFMA with float32 accumulation) that maps
 * an observation and its kernel widths to an estimate of the 200-iteration
 * RL result, used to seed a short RL run (ClarityEx_RLPlanReset).  Volume
 * dims must be multiples of 4 and an x-y plane under 2^32 / 96 voxels
 * (CLARITY_INVALID_ARGUMENT otherwise).  Parameters are float32
 * (He-initialised from `seed`). */
typedef struct ClarityEx_UNet_t *ClarityEx_UNet;

C_FUNC_DEF ClarityResult_t ClarityEx_UNetCreate(ClarityEx_UNet *net, unsigned seed);
C_FUNC_DEF ClarityResult_t ClarityEx_UNetDestroy(ClarityEx_UNet net);
C_FUNC_DEF int64_t ClarityEx_UNetNumParams(ClarityEx_UNet net);
C_FUNC_DEF ClarityResult_t ClarityEx_UNetGetParams(ClarityEx_UNet net, float *params);
C_FUNC_DEF ClarityResult_t ClarityEx_UNetSetParams(ClarityEx_UNet net,
                                                   const float *params);

/* out = the network's estimate for `obs` (dims floats), in obs units. */
C_FUNC_DEF ClarityResult_t ClarityEx_UNetPredict(ClarityEx_UNet net, const float *obs,
                                                 Clarity_Dim3 dims, float sigmaXY,
                                                 float sigmaZ, float *out,
                                                 int onDevice);

/* One training step on (obs, target): mean squared error of the estimate
 * against target, both divided by mean(obs); Adam (beta 0.9 / 0.999) with
 * `learningRate` (0 = gradients only, no update).  `loss` (may be NULL)
 * receives the loss, `grads` (host, NumParams floats, may be NULL) the
 * gradient of every parameter. */
C_FUNC_DEF ClarityResult_t ClarityEx_UNetTrainStep(
    ClarityEx_UNet net, const float *obs, const float *target, Clarity_Dim3 dims,
    float sigmaXY, float sigmaZ, float learningRate, float *loss, float *grads,
    int onDevice);

/* Clarity_DivideArraysComponentWise (reference include/internal/
 * ComputePrimitives.h, src/ComputePrimitives.cxx:121-158) over n host floats,
 * computed as the fused Richardson-Lucy passes compute their ratio:
 * result = value where |b| < 1e-5 (the reference's double compare, exactly),
 * else a / b by the hardware reciprocal and one Newton step, within 1 ulp of
 * IEEE division for 1e-5 <= |b| < 2^126. */
C_FUNC_DEF ClarityResult_t ClarityEx_DivideArraysComponentWise(float *result,
                                                               const float *a,
                                                               const float *b,
                                                               float value,
                                                               int64_t n);

/* Clarity_Convolve on DEVICE buffers with no allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * inImage (imageDim floats), kernelImage (kernelDim floats) and outImage
 * (imageDim floats) are device pointers; the result is Clarity_Convolve's
 * (reference work domain image + kernel per axis, generic FFT passes).
 * Enqueued on `stream` (a hipStream_t; NULL = the library stream) with no
 * host synchronisation: synchronise that stream before reading outImage.
 * Every intermediate lives in one static per-device scratch array of 32 MiB
 * (held by each GPU the library runs on until the process ends): the work
 * domain (imageDim + kernelDim per axis, each axis <= 4096 with prime
 * factors <= 64) must satisfy 4 V + 16 C + 8 (nx + ny + nz) <= 32 MiB
 * (V = nx ny nz, C = nz ny (nx/2 + 1), each term rounded up to 256 bytes),
 * e.g. a 128x128x64 image with a 32x32x32 kernel (work 160x160x96); larger
 * shapes return CLARITY_INVALID_ARGUMENT, other axis lengths
 * CLARITY_FFT_FAILED, before anything is enqueued.  Calls share that array:
 * calls that overlap on one GPU (other streams, other threads' streams)
 * race, and the array keeps the last call's contents. */
C_FUNC_DEF ClarityResult_t ClarityEx_ConvolveNoAlloc(const float *inImage,
                                                     Clarity_Dim3 imageDim,
                                                     const float *kernelImage,
                                                     Clarity_Dim3 kernelDim,
                                                     float *outImage, void *stream);

/* Engine of the last Convolve, Wiener, Jansen-van Cittert, (blind) ML or
 * blind RL call: "fast" (the fused passes) or "generic" ("" before any). */
C_FUNC_DEF const char *ClarityEx_LastEngine();

/* Blocks until all work queued on the library stream has finished. */
C_FUNC_DEF ClarityResult_t ClarityEx_Synchronize();

/* Name of the GPU in use ("" before Clarity_Register). */
C_FUNC_DEF const char *ClarityEx_DeviceName();

/* Test hook: from now on every `every`-th device allocation fails with
 * CLARITY_DEVICE_OUT_OF_MEMORY (1 = all of them, 0 = off).  The environment
 * variable CLARITY_FAULT_DEVICE_OOM sets the same at Clarity_Register.  Lets
 * callers check their out-of-memory handling without exhausting HBM. */
C_FUNC_DEF ClarityResult_t ClarityEx_InjectDeviceOOM(int64_t every);

#endif /* __CLARITY_EX_H_ */
