// Engine interface behind ClarityEx_RLPlan and Clarity_MaximumLikelihood*.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "work.hpp"

namespace clarity {

class RLEngine {
 public:
  virtual ~RLEngine() = default;
  virtual ClarityResult_t allocate() = 0;
  // obs_d: imageDim floats, psf_d: kernelDim floats (device).  Computes the
  // OTF and E_obs; leaves the estimate undefined until reset().
  virtual ClarityResult_t set_inputs(const float *obs_d, const float *psf_d,
                                     hipStream_t s) = 0;
  // init_d: imageDim floats on the device, or nullptr for the observation.
  virtual ClarityResult_t reset(const float *init_d, hipStream_t s) = 0;
  virtual ClarityResult_t iterate(int n, hipStream_t s) = 0;
  // Writes the energy-scaled estimate (imageDim floats) to out_d.
  virtual ClarityResult_t estimate(float *out_d, hipStream_t s) = 0;
  // Device address of the last rescale factor.
  virtual const float *scale_ptr() const = 0;
  virtual const char *name() const = 0;
  // The FFT work domain the engine transforms over.
  Dims work_dims() const { return w_.wd; }
  virtual int describe(const char **names, double *bytes, int *launches,
                       int cap) {
    (void)names; (void)bytes; (void)launches; (void)cap;
    return 0;
  }
  virtual ClarityResult_t set_timing(bool on) {
    (void)on;
    return CLARITY_INVALID_OPERATION;
  }
  virtual int timings(int *count, double *ms, int cap, bool reset) {
    (void)count; (void)ms; (void)cap; (void)reset;
    return 0;
  }
  // Jansen-van Cittert engines (make_fast_jvc): the update's A and 1/A^2,
  // after set_inputs.
  virtual ClarityResult_t set_jvc_constants(float A, float invASq, hipStream_t s) {
    (void)A; (void)invASq; (void)s;
    return CLARITY_INVALID_OPERATION;
  }
  // Picks launch shapes by timing candidates on the current inputs (the
  // estimate is saved and restored); writes a description of the choice to
  // `desc` (cap bytes).  Engines without tunable shapes do nothing.
  virtual ClarityResult_t autotune(int iterations, char *desc, int cap, hipStream_t s) {
    (void)iterations; (void)s;
    if (desc && cap > 0) desc[0] = 0;
    return CLARITY_SUCCESS;
  }

 protected:
  Work w_;
};

std::unique_ptr<RLEngine> make_generic_rl(const Work &geom);
// nullptr when the shape is not one the fused pipeline handles.
std::unique_ptr<RLEngine> make_fast_rl(const Work &geom);
// The kernel half-step of alternating blind RL on the fused pipeline
// (estimate = the kernel, blurring input = the image estimate; same even
// image and kernel dimensions, reference work domain); nullptr otherwise.
std::unique_ptr<RLEngine> make_fast_rl_kernel_step(const Work &geom);

// Jansen-van Cittert iterations on the fused passes over the reference work
// domain (its "image" is the padded observation, its estimate the whole
// domain); nullptr when the domain is not one of the fused engine's sizes.
std::unique_ptr<RLEngine> make_fast_jvc(const Work &geom);

// RL without the energy rescale on the fused engine (the estimate's scale
// stays 1); nullptr when the shape is not the fused engine's.
std::unique_ptr<RLEngine> make_fast_rl_unscaled(const Work &geom);

// A single convolution (alias-free domain) or Wiener filter (reference
// domain) on the fused passes: reset(input), iterate(1), estimate(out);
// nullptr when the shape is not the fused engine's.
std::unique_ptr<RLEngine> make_fast_convolve(const Work &geom);
std::unique_ptr<RLEngine> make_fast_wiener(const Work &geom, float eps);

// RL iterations of several engines together (for each pass, one launch per
// kernel instantiation over all the volumes that use it).
class RLGroup {
 public:
  virtual ~RLGroup() = default;
  virtual ClarityResult_t iterate(int n, hipStream_t s) = 0;
};
// nullptr unless every engine is a fused one (allocated, inputs set).
std::unique_ptr<RLGroup> make_rl_group(const std::vector<RLEngine *> &engines);

}  // namespace clarity
