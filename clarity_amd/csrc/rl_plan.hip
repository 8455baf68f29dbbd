// Richardson-Lucy ("maximum likelihood") driver and the resident RL plan.
//
// reference src/MaximumLikelihoodDeconvolve.cxx:48-81 (update), :84-153
// (driver), :255-332 (public wrapper).  Per iteration:
//   s1 = H (*) g;  r = guard(obs / s1, 0);  s2 = H (*) r;  g' = g * s2;
//   g  = g' * (E_obs / sum(g'))
// with the plain OTF H in both convolutions.  The plan keeps the padded
// observation, the OTF and the estimate in HBM; the rescale factor lives in
// device memory, so iterations queue back to back without host syncs.
//
// Two engines implement the plan: the generic one below (any supported
// shape, FFTW layout, one kernel per step) and the fused pruned pipeline in
// rl_fast.hip for power-of-two work shapes.
#include "common.hpp"
#include "rl_plan.hpp"
#include "work.hpp"

namespace clarity {

namespace {

class GenericRL : public RLEngine {
 public:
  explicit GenericRL(const Work &geom) { w_.set_geometry(geom); }

  ClarityResult_t allocate() override {
    CL_TRY(w_.obs.alloc(sizeof(float) * w_.V));
    CL_TRY(w_.psf.alloc(sizeof(float) * w_.V));
    CL_TRY(otf_.alloc(sizeof(float2) * w_.C));
    CL_TRY(spec_.alloc(sizeof(float2) * w_.C));
    CL_TRY(g_.alloc(sizeof(float) * w_.V));
    CL_TRY(s1_.alloc(sizeof(float) * w_.V));
    CL_TRY(s2_.alloc(sizeof(float) * w_.V));
    CL_TRY(parts_.alloc(sizeof(double) * kParts));
    CL_TRY(scal_.alloc(sizeof(float) * 4));
    return CLARITY_SUCCESS;
  }

  ClarityResult_t set_inputs(const float *obs_d, const float *psf_d,
                             hipStream_t s) override {
    const int zero[3] = {0, 0, 0};
    const int kshift[3] = {-w_.kd.x / 2, -w_.kd.y / 2, -w_.kd.z / 2};
    CL_TRY(k_pad_shift(w_.obs.f(), w_.wd, obs_d, w_.id, zero, 0.f, s));
    CL_TRY(k_pad_shift(w_.psf.f(), w_.wd, psf_d, w_.kd, kshift, 0.f, s));
    CL_TRY(fft3d_r2c(w_.wd.x, w_.wd.y, w_.wd.z, w_.psf.f(), otf_.c(), s));
    // E_obs = sum(padded observation) (reference :131-133)
    CL_TRY(k_sum_partials(w_.obs.f(), w_.V, parts_.d(), kParts, s));
    CL_TRY(k_finish_sum(parts_.d(), kParts, scal_.f(), nullptr, nullptr, s));
    return CLARITY_SUCCESS;
  }

  ClarityResult_t reset(const float *init_d, hipStream_t s) override {
    if (init_d) {
      const int zero[3] = {0, 0, 0};
      CL_TRY(k_pad_shift(g_.f(), w_.wd, init_d, w_.id, zero, 0.f, s));
    } else {
      CL_HIP(hipMemcpyAsync(g_.p, w_.obs.p, sizeof(float) * w_.V,
                            hipMemcpyDeviceToDevice, s));
    }
    CL_TRY(k_fill(scal_.f() + 2, 1, 1.0f, s));
    return CLARITY_SUCCESS;
  }

  ClarityResult_t iterate(int n, hipStream_t s) override {
    float *E = scal_.f(), *S = scal_.f() + 1, *c = scal_.f() + 2;
    for (int k = 0; k < n; k++) {
      CL_TRY(convolve_otf(w_, g_.f(), otf_.c(), s1_.f(), spec_.c(), s));
      CL_TRY(k_divide(s1_.f(), w_.obs.f(), s1_.f(), 0.f, w_.V, s));
      CL_TRY(convolve_otf(w_, s1_.f(), otf_.c(), s2_.f(), spec_.c(), s));
      CL_TRY(k_multiply_sum(g_.f(), g_.f(), s2_.f(), w_.V, parts_.d(), kParts, s));
      CL_TRY(k_finish_sum(parts_.d(), kParts, S, E, c, s));
      CL_TRY(k_scale_dev(g_.f(), g_.f(), w_.V, c, s));
    }
    return CLARITY_SUCCESS;
  }

  ClarityResult_t estimate(float *out_d, hipStream_t s) override {
    return k_clip(out_d, w_.id, g_.f(), w_.wd, s);
  }

  const float *scale_ptr() const override { return scal_.f() + 2; }
  const char *name() const override { return "generic"; }

 private:
  DevBuf otf_, spec_, g_, s1_, s2_, parts_, scal_;
};

}  // namespace

std::unique_ptr<RLEngine> make_generic_rl(const Work &geom) {
  return std::unique_ptr<RLEngine>(new GenericRL(geom));
}

}  // namespace clarity

using namespace clarity;

struct ClarityEx_RLPlan_t {
  RegHold reg;  // first: released after the engine's buffers
  Work geom;
  std::unique_ptr<RLEngine> eng;
  bool have_inputs = false;
};

namespace {

ClarityResult_t plan_create(ClarityEx_RLPlan *plan, Clarity_Dim3 id,
                            Clarity_Dim3 kd) {
  if (!plan) return CLARITY_INVALID_ARGUMENT;
  *plan = nullptr;
  float dummy = 0.f;
  std::unique_ptr<ClarityEx_RLPlan_t> p(new ClarityEx_RLPlan_t);
  CL_TRY(work_geometry(p->geom, &dummy, id, &dummy, kd, &dummy));
  CL_TRY(p->reg.acquire());
  p->eng = make_fast_rl(p->geom);
  if (!p->eng || std::getenv("CLARITY_FORCE_GENERIC"))
    p->eng = make_generic_rl(p->geom);
  CL_TRY(p->eng->allocate());
  *plan = p.release();
  return CLARITY_SUCCESS;
}

ClarityResult_t plan_set_inputs(ClarityEx_RLPlan p, const float *in,
                                const float *kernel, int on_device) {
  if (!p || !in || !kernel) return CLARITY_INVALID_ARGUMENT;
  hipStream_t s = rt().stream;
  const Work &w = p->geom;
  if (on_device) {
    CL_TRY(p->eng->set_inputs(in, kernel, s));
  } else {
    DevBuf a, b;
    CL_TRY(a.alloc(sizeof(float) * w.IV));
    CL_TRY(b.alloc(sizeof(float) * w.KV));
    CL_TRY(host_to_device(a.p, in, sizeof(float) * w.IV, s));
    CL_TRY(host_to_device(b.p, kernel, sizeof(float) * w.KV, s));
    CL_TRY(p->eng->set_inputs(a.f(), b.f(), s));
    CL_HIP(hipStreamSynchronize(s));
  }
  p->have_inputs = true;
  return p->eng->reset(nullptr, s);
}

ClarityResult_t plan_reset(ClarityEx_RLPlan p, const float *init, int on_device) {
  if (!p) return CLARITY_INVALID_ARGUMENT;
  if (!p->have_inputs) return CLARITY_INVALID_OPERATION;
  hipStream_t s = rt().stream;
  if (!init || on_device) return p->eng->reset(init, s);
  DevBuf a;
  CL_TRY(a.alloc(sizeof(float) * p->geom.IV));
  CL_TRY(host_to_device(a.p, init, sizeof(float) * p->geom.IV, s));
  CL_TRY(p->eng->reset(a.f(), s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

ClarityResult_t plan_get(ClarityEx_RLPlan p, float *out, int on_device) {
  if (!p || !out) return CLARITY_INVALID_ARGUMENT;
  if (!p->have_inputs) return CLARITY_INVALID_OPERATION;
  hipStream_t s = rt().stream;
  if (on_device) {
    CL_TRY(p->eng->estimate(out, s));
  } else {
    DevBuf a;
    CL_TRY(a.alloc(sizeof(float) * p->geom.IV));
    CL_TRY(p->eng->estimate(a.f(), s));
    CL_TRY(device_to_host(out, a.p, sizeof(float) * p->geom.IV, s));
  }
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

// reference src/MaximumLikelihoodDeconvolve.cxx:255-332
ClarityResult_t ml_call(float *in, Clarity_Dim3 id, float *kernel,
                        Clarity_Dim3 kd, float *init, float *out,
                        int iterations) {
  if (!in || !kernel || !out) return CLARITY_INVALID_ARGUMENT;
  if (iterations < 0) return CLARITY_INVALID_ARGUMENT;
  ClarityEx_RLPlan p = nullptr;
  CL_TRY(plan_create(&p, id, kd));
  std::unique_ptr<ClarityEx_RLPlan_t> hold(p);
  rt().last_engine = p->eng->name();
  CL_TRY(plan_set_inputs(p, in, kernel, 0));
  if (init) CL_TRY(plan_reset(p, init, 0));
  CL_TRY(p->eng->iterate(iterations, rt().stream));
  return plan_get(p, out, 0);
}

// Independent volumes on a pool of streams: volume v runs on stream
// v % nstreams, so up to nstreams volumes' kernels overlap on the GPU.  Each
// volume is computed exactly as ml_call would compute it alone.
ClarityResult_t ml_batch(int count, float *const *ins, const Clarity_Dim3 *ids,
                         float *const *kers, const Clarity_Dim3 *kds,
                         float *const *outs, int iterations) {
  if (count < 0 || iterations < 0) return CLARITY_INVALID_ARGUMENT;
  if (count == 0) return CLARITY_SUCCESS;
  if (!ins || !ids || !kers || !kds || !outs) return CLARITY_INVALID_ARGUMENT;
  for (int v = 0; v < count; v++)
    if (!ins[v] || !kers[v] || !outs[v] || !dims_ok(ids[v]) || !dims_ok(kds[v]))
      return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  struct Vol {
    std::unique_ptr<ClarityEx_RLPlan_t> plan;
    DevBuf a, b;
  };
  const int nstreams = count < 8 ? count : 8;
  std::vector<hipStream_t> streams(nstreams, nullptr);
  ClarityResult_t rc = CLARITY_SUCCESS;
  for (int i = 0; i < nstreams && rc == CLARITY_SUCCESS; i++)
    rc = from_hip(hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking));
  std::vector<Vol> vols(count);
  for (int v = 0; v < count && rc == CLARITY_SUCCESS; v++) {
    hipStream_t s = streams[v % nstreams];
    ClarityEx_RLPlan p = nullptr;
    rc = plan_create(&p, ids[v], kds[v]);
    if (rc != CLARITY_SUCCESS) break;
    vols[v].plan.reset(p);
    const Work &w = p->geom;
    rc = vols[v].a.alloc(sizeof(float) * w.IV);
    if (rc == CLARITY_SUCCESS) rc = vols[v].b.alloc(sizeof(float) * w.KV);
    if (rc == CLARITY_SUCCESS)
      rc = host_to_device(vols[v].a.p, ins[v], sizeof(float) * w.IV, s);
    if (rc == CLARITY_SUCCESS)
      rc = host_to_device(vols[v].b.p, kers[v], sizeof(float) * w.KV, s);
    if (rc == CLARITY_SUCCESS) rc = p->eng->set_inputs(vols[v].a.f(), vols[v].b.f(), s);
    if (rc == CLARITY_SUCCESS) rc = p->eng->reset(nullptr, s);
    p->have_inputs = true;
  }
  // All fused: one launch per pass and kernel instantiation over every
  // volume (make_rl_group), on the first stream once the set-up is done.
  // Otherwise (or with CLARITY_BATCH_STREAMS set) each volume on its own
  // plan, interleaved in rounds of a few iterations over the streams, so
  // all stay busy until the end; within a round a volume's iterations
  // chain and replay its captured graphs.
  std::unique_ptr<RLGroup> group;
  if (rc == CLARITY_SUCCESS && count > 1 && !std::getenv("CLARITY_BATCH_STREAMS")) {
    std::vector<RLEngine *> engines;
    for (int v = 0; v < count; v++) engines.push_back(vols[v].plan->eng.get());
    group = make_rl_group(engines);
  }
  if (group) {
    for (int i = 0; i < nstreams && rc == CLARITY_SUCCESS; i++)
      rc = from_hip(hipStreamSynchronize(streams[i]));
    if (rc == CLARITY_SUCCESS) rc = group->iterate(iterations, streams[0]);
    if (rc == CLARITY_SUCCESS) rc = from_hip(hipStreamSynchronize(streams[0]));
  } else {
    constexpr int kRound = 10;
    for (int done = 0; done < iterations && rc == CLARITY_SUCCESS; done += kRound) {
      const int n = iterations - done < kRound ? iterations - done : kRound;
      for (int v = 0; v < count && rc == CLARITY_SUCCESS; v++)
        rc = vols[v].plan->eng->iterate(n, streams[v % nstreams]);
    }
  }
  for (int v = 0; v < count && rc == CLARITY_SUCCESS; v++) {
    hipStream_t s = streams[v % nstreams];
    const Work &w = vols[v].plan->geom;
    rc = vols[v].plan->eng->estimate(vols[v].a.f(), s);
    if (rc == CLARITY_SUCCESS)
      rc = device_to_host(outs[v], vols[v].a.p, sizeof(float) * w.IV, s);
  }
  for (hipStream_t s : streams) {
    if (!s) continue;
    ClarityResult_t r2 = from_hip(hipStreamSynchronize(s));
    if (rc == CLARITY_SUCCESS) rc = r2;
  }
  group.reset();  // its graphs refer to the volumes' buffers
  vols.clear();   // frees device buffers while the streams still exist
  for (hipStream_t s : streams)
    if (s) (void)hipStreamDestroy(s);
  return rc;
}

}  // namespace

#define API_LOCK std::lock_guard<std::recursive_mutex> lk_(rt().mu)

extern "C" {

ClarityResult_t Clarity_MaximumLikelihoodDeconvolve(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *outImage, int iterations) {
  API_LOCK;
  return ml_call(inImage, imageDim, kernelImage, kernelDim, nullptr, outImage,
                 iterations);
}

ClarityResult_t ClarityEx_MaximumLikelihoodDeconvolveFrom(
    float *inImage, Clarity_Dim3 imageDim, float *kernelImage,
    Clarity_Dim3 kernelDim, float *initialImage, float *outImage,
    int iterations) {
  API_LOCK;
  return ml_call(inImage, imageDim, kernelImage, kernelDim, initialImage,
                 outImage, iterations);
}

ClarityResult_t ClarityEx_MaximumLikelihoodDeconvolveBatch(
    int count, float *const *inImages, const Clarity_Dim3 *imageDims,
    float *const *kernelImages, const Clarity_Dim3 *kernelDims,
    float *const *outImages, int iterations) {
  API_LOCK;
  return ml_batch(count, inImages, imageDims, kernelImages, kernelDims,
                  outImages, iterations);
}

ClarityResult_t ClarityEx_RLPlanCreate(ClarityEx_RLPlan *plan,
                                       Clarity_Dim3 imageDim,
                                       Clarity_Dim3 kernelDim) {
  API_LOCK;
  return plan_create(plan, imageDim, kernelDim);
}

ClarityResult_t ClarityEx_RLPlanSetInputs(ClarityEx_RLPlan plan,
                                          const float *inImage,
                                          const float *kernelImage,
                                          int onDevice) {
  API_LOCK;
  return plan_set_inputs(plan, inImage, kernelImage, onDevice);
}

ClarityResult_t ClarityEx_RLPlanReset(ClarityEx_RLPlan plan,
                                      const float *initial, int onDevice) {
  API_LOCK;
  return plan_reset(plan, initial, onDevice);
}

ClarityResult_t ClarityEx_RLPlanIterate(ClarityEx_RLPlan plan, int iterations) {
  API_LOCK;
  if (!plan || iterations < 0) return CLARITY_INVALID_ARGUMENT;
  if (!plan->have_inputs) return CLARITY_INVALID_OPERATION;
  return plan->eng->iterate(iterations, rt().stream);
}

ClarityResult_t ClarityEx_RLPlanGetEstimate(ClarityEx_RLPlan plan, float *out,
                                            int onDevice) {
  API_LOCK;
  return plan_get(plan, out, onDevice);
}

ClarityResult_t ClarityEx_RLPlanGetScale(ClarityEx_RLPlan plan, float *scale) {
  API_LOCK;
  if (!plan || !scale) return CLARITY_INVALID_ARGUMENT;
  hipStream_t s = rt().stream;
  CL_TRY(device_to_host(scale, plan->eng->scale_ptr(), sizeof(float), s));
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

ClarityResult_t ClarityEx_RLPlanDestroy(ClarityEx_RLPlan plan) {
  API_LOCK;
  if (!plan) return CLARITY_INVALID_ARGUMENT;
  if (rt().stream) (void)hipStreamSynchronize(rt().stream);
  delete plan;
  return CLARITY_SUCCESS;
}

int ClarityEx_RLPlanDescribe(ClarityEx_RLPlan plan, const char **names,
                             double *bytes, int *launches, int cap) {
  API_LOCK;
  if (!plan) return 0;
  return plan->eng->describe(names, bytes, launches, cap);
}

ClarityResult_t ClarityEx_RLPlanSetTiming(ClarityEx_RLPlan plan, int enable) {
  API_LOCK;
  if (!plan) return CLARITY_INVALID_ARGUMENT;
  return plan->eng->set_timing(enable != 0);
}

ClarityResult_t ClarityEx_RLPlanAutotune(ClarityEx_RLPlan plan, int iterations,
                                         char *description, int cap) {
  API_LOCK;
  if (!plan || iterations < 1 || (description && cap < 0)) return CLARITY_INVALID_ARGUMENT;
  if (!plan->have_inputs) return CLARITY_INVALID_OPERATION;
  return plan->eng->autotune(iterations, description, cap, rt().stream);
}

int ClarityEx_RLPlanTimings(ClarityEx_RLPlan plan, int *count, double *ms,
                            int cap, int reset) {
  API_LOCK;
  if (!plan) return 0;
  return plan->eng->timings(count, ms, cap, reset != 0);
}

const char *ClarityEx_RLPlanEngine(ClarityEx_RLPlan plan) {
  return plan ? plan->eng->name() : "";
}

ClarityResult_t ClarityEx_RLPlanWorkDims(ClarityEx_RLPlan plan, Clarity_Dim3 *workDim) {
  if (!plan || !workDim) return CLARITY_INVALID_ARGUMENT;
  const Dims d = plan->eng->work_dims();
  workDim->x = d.x;
  workDim->y = d.y;
  workDim->z = d.z;
  return CLARITY_SUCCESS;
}

}  // extern "C"
