// Library runtime: registration (reference src/Clarity.cxx:41-150), device
// selection, the library stream and the device-memory cache.
#include "common.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace clarity {

// Scratch of the allocation-free entry point (ClarityEx_ConvolveNoAlloc):
// every GPU the library runs on holds its own copy of this array until the
// process ends.  Host code never passes the array's name to a kernel (on
// the host it only stands for the symbol); it uses the device address that
// Clarity_Register looks up with hipGetSymbolAddress, once per device, and
// keeps in Runtime::noalloc_scratch.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

Runtime &rt() {
  static Runtime r;
  return r;
}

// Entry points called before Clarity_Register initialise the device on
// first use: the reference's functions work unregistered (src/Clarity.cxx:
// 20-55 only sets up threads and the device flag), and its own raw-image
// app never registers (apps/DeconvolveRawImages/DeconvolveRawImages.cxx:
// 98-100).  The implicit registration counts as one client and holds the
// device until a Clarity_UnRegister or process exit.
bool ready() {
  Runtime &r = rt();
  std::lock_guard<std::recursive_mutex> lk(r.mu);
  if (r.refcount > 0 && r.device >= 0) return true;
  return Clarity_Register() == CLARITY_SUCCESS;
}

ClarityResult_t from_hip(hipError_t e) {
  if (e == hipSuccess) return CLARITY_SUCCESS;
  if (e == hipErrorOutOfMemory) return CLARITY_DEVICE_OUT_OF_MEMORY;
  if (std::getenv("CLARITY_LOG"))
    std::fprintf(stderr, "clarity: HIP error %d (%s)\n", (int)e,
                 hipGetErrorString(e));
  return CLARITY_INVALID_OPERATION;
}

ClarityResult_t launch_check() {
  hipError_t e = hipGetLastError();
  return from_hip(e);
}

ClarityResult_t dev_alloc(void **p, size_t bytes) {
  Runtime &r = rt();
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  r.alloc_count++;
  // Fault injection (ClarityEx_InjectDeviceOOM, or CLARITY_FAULT_DEVICE_OOM
  // read at Register): every k-th device allocation fails as if HBM were
  // exhausted.
  if (r.fault_every > 0 && r.alloc_count % r.fault_every == 0)
    return CLARITY_DEVICE_OUT_OF_MEMORY;
  auto it = r.cache.find(bytes);
  if (it != r.cache.end()) {
    *p = it->second;
    r.cache_bytes -= it->first;
    r.cache.erase(it);
    return CLARITY_SUCCESS;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // give cached blocks back and retry once
    dev_cache_clear();
    e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return CLARITY_DEVICE_OUT_OF_MEMORY;
    }
  }
  return CLARITY_SUCCESS;
}

namespace {

constexpr size_t kStageBytes = (size_t)32 << 20;  // per staging buffer
constexpr size_t kStageMin = (size_t)4 << 20;     // smaller copies go directly

bool stage_ready(Runtime &r) {
  if (r.stage[0] && r.stage[1]) return true;
  for (int i = 0; i < 2; i++) {
    if (!r.stage[i] && hipHostMalloc(&r.stage[i], kStageBytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      r.stage[i] = nullptr;
      return false;
    }
    if (!r.stage_ev[i] &&
        hipEventCreateWithFlags(&r.stage_ev[i], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      r.stage_ev[i] = nullptr;
      return false;
    }
  }
  return true;
}

// memcpy split over up to 8 threads (64-byte aligned pieces)
void par_copy(void *dst, const void *src, size_t n) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int T = (int)std::min<unsigned>(8u, hw ? hw : 1u);
  if (T <= 1 || n < ((size_t)1 << 20)) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t per = ((n + T - 1) / T + 63) & ~(size_t)63;
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++) {
    const size_t o = (size_t)t * per;
    if (o >= n) break;
    const size_t len = std::min(per, n - o);
    th.emplace_back([=] {
      std::memcpy(static_cast<char *>(dst) + o, static_cast<const char *>(src) + o, len);
    });
  }
  for (auto &x : th) x.join();
}

}  // namespace

ClarityResult_t host_to_device(void *dst, const void *src, size_t bytes, hipStream_t s) {
  Runtime &r = rt();
  if (bytes < kStageMin || !stage_ready(r))
    return from_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  for (size_t o = 0; o < bytes; o += kStageBytes) {
    const int b = r.stage_next;
    r.stage_next ^= 1;
    const size_t len = std::min(kStageBytes, bytes - o);
    CL_HIP(hipEventSynchronize(r.stage_ev[b]));  // its previous DMA is done
    par_copy(r.stage[b], static_cast<const char *>(src) + o, len);
    CL_HIP(hipMemcpyAsync(static_cast<char *>(dst) + o, r.stage[b], len,
                          hipMemcpyHostToDevice, s));
    CL_HIP(hipEventRecord(r.stage_ev[b], s));
  }
  return CLARITY_SUCCESS;
}

ClarityResult_t device_to_host(void *dst, const void *src, size_t bytes, hipStream_t s) {
  Runtime &r = rt();
  if (bytes < kStageMin || !stage_ready(r)) {
    CL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    CL_HIP(hipStreamSynchronize(s));
    return CLARITY_SUCCESS;
  }
  const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
  const int base = r.stage_next;
  r.stage_next = (int)((base + nch) & 1);
  // chunk i goes through buffer (base + i) % 2; issuing it waits for that
  // buffer's previous DMA (chunk i - 2's host copy has finished already)
  const auto issue = [&](size_t i) -> ClarityResult_t {
    const int b = (int)((base + i) & 1);
    const size_t o = i * kStageBytes, len = std::min(kStageBytes, bytes - o);
    CL_HIP(hipEventSynchronize(r.stage_ev[b]));
    CL_HIP(hipMemcpyAsync(r.stage[b], static_cast<const char *>(src) + o, len,
                          hipMemcpyDeviceToHost, s));
    CL_HIP(hipEventRecord(r.stage_ev[b], s));
    return CLARITY_SUCCESS;
  };
  CL_TRY(issue(0));
  for (size_t i = 0; i < nch; i++) {
    if (i + 1 < nch) CL_TRY(issue(i + 1));
    const int b = (int)((base + i) & 1);
    const size_t o = i * kStageBytes, len = std::min(kStageBytes, bytes - o);
    CL_HIP(hipEventSynchronize(r.stage_ev[b]));
    par_copy(static_cast<char *>(dst) + o, r.stage[b], len);
  }
  return CLARITY_SUCCESS;
}

void dev_release(void *p, size_t bytes) {
  if (!p) return;
  Runtime &r = rt();
  if (bytes == 0) bytes = 256;
  // Keep freed blocks for reuse, up to a bound well inside 288 GB of HBM.
  const size_t kCacheLimit = (size_t)64 << 30;
  if (r.cache_bytes + bytes <= kCacheLimit) {
    r.cache.emplace(bytes, p);
    r.cache_bytes += bytes;
  } else {
    (void)hipFree(p);
  }
}

void dev_cache_clear() {
  Runtime &r = rt();
  if (r.stream) (void)hipStreamSynchronize(r.stream);
  for (auto &kv : r.cache) (void)hipFree(kv.second);
  r.cache.clear();
  r.cache_bytes = 0;
}

}  // namespace clarity

using namespace clarity;

extern "C" {

Clarity_Dim3 Clarity_Dim3FromArray(int dimArray[3]) {
  Clarity_Dim3 d;
  d.x = dimArray[0];
  d.y = dimArray[1];
  d.z = dimArray[2];
  return d;
}

ClarityResult_t Clarity_Register() {
  Runtime &r = rt();
  std::lock_guard<std::recursive_mutex> lk(r.mu);
  if (r.refcount > 0) {
    r.refcount++;
    return CLARITY_SUCCESS;
  }
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count < 1) {
    (void)hipGetLastError();
    std::fprintf(stderr, "clarity: no HIP device available\n");
    return CLARITY_INVALID_OPERATION;
  }
  int dev = 0;
  if (const char *s = std::getenv("CLARITY_DEVICE")) dev = std::atoi(s);
  if (dev < 0 || dev >= count) return CLARITY_INVALID_ARGUMENT;
  CL_HIP(hipSetDevice(dev));
  // the scratch array's address on this device (the device is current);
  // a failed lookup leaves nullptr, and only the allocation-free entry
  // point, which then returns CLARITY_INVALID_OPERATION, depends on it
  if (r.noalloc_scratch.find(dev) == r.noalloc_scratch.end()) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      ptr = nullptr;
    }
    r.noalloc_scratch[dev] = ptr;
  }
  hipDeviceProp_t prop;
  CL_HIP(hipGetDeviceProperties(&prop, dev));
  CL_HIP(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
  r.device = dev;
  // the marketing name can be empty; the architecture name never is
  r.device_name = prop.name[0] ? std::string(prop.name) + " (" + prop.gcnArchName + ")"
                               : std::string(prop.gcnArchName);
  r.alloc_count = 0;
  r.fault_every = 0;
  if (const char *f = std::getenv("CLARITY_FAULT_DEVICE_OOM"))
    r.fault_every = std::atoll(f);
  if (std::getenv("CLARITY_LOG"))
    std::fprintf(stderr, "clarity: device %d '%s' (%s), %zu MB\n", dev,
                 prop.name, prop.gcnArchName,
                 (size_t)(prop.totalGlobalMem >> 20));
  r.refcount = 1;
  return CLARITY_SUCCESS;
}

ClarityResult_t Clarity_UnRegister() {
  Runtime &r = rt();
  std::lock_guard<std::recursive_mutex> lk(r.mu);
  if (r.refcount <= 0) return CLARITY_SUCCESS;
  if (--r.refcount > 0) return CLARITY_SUCCESS;
  if (r.stream) (void)hipStreamSynchronize(r.stream);
  for (auto &kv : r.twiddles) (void)hipFree(kv.second);
  r.twiddles.clear();
  for (auto &kv : r.tables) (void)hipFree(kv.second);
  r.tables.clear();
  for (int i = 0; i < 2; i++) {
    if (r.stage[i]) (void)hipHostFree(r.stage[i]);
    if (r.stage_ev[i]) (void)hipEventDestroy(r.stage_ev[i]);
    r.stage[i] = nullptr;
    r.stage_ev[i] = nullptr;
  }
  r.stage_next = 0;
  dev_cache_clear();
  if (r.stream) (void)hipStreamDestroy(r.stream);
  r.stream = nullptr;
  r.device = -1;
  return CLARITY_SUCCESS;
}

ClarityResult_t Clarity_Init() { return Clarity_Register(); }
ClarityResult_t Clarity_Shutdown() { return Clarity_UnRegister(); }

ClarityResult_t Clarity_SetNumberOfThreads(unsigned n) {
  rt().host_threads = n;
  return CLARITY_SUCCESS;
}

char *Clarity_ResultString(ClarityResult_t result) {
  static char s_fft[] = "CLARITY_FFT_FAILED";
  static char s_oom[] = "CLARITY_OUT_OF_MEMORY";
  static char s_doom[] = "CLARITY_DEVICE_OUT_OF_MEMORY";
  static char s_iop[] = "CLARITY_INVALID_OPERATION";
  static char s_iarg[] = "CLARITY_INVALID_ARGUMENT";
  static char s_ok[] = "CLARITY_SUCCESS";
  static char s_unk[] = "<unknown Clarity result>";
  switch (result) {
    case CLARITY_FFT_FAILED: return s_fft;
    case CLARITY_OUT_OF_MEMORY: return s_oom;
    case CLARITY_DEVICE_OUT_OF_MEMORY: return s_doom;
    case CLARITY_INVALID_OPERATION: return s_iop;
    case CLARITY_INVALID_ARGUMENT: return s_iarg;
    case CLARITY_SUCCESS: return s_ok;
    default: return s_unk;
  }
}

ClarityResult_t ClarityEx_Synchronize() {
  if (!ready()) return CLARITY_INVALID_OPERATION;
  CL_HIP(hipStreamSynchronize(rt().stream));
  return CLARITY_SUCCESS;
}

const char *ClarityEx_DeviceName() { return rt().device_name.c_str(); }

const char *ClarityEx_LastEngine() { return rt().last_engine; }

ClarityResult_t ClarityEx_InjectDeviceOOM(int64_t every) {
  Runtime &r = rt();
  std::lock_guard<std::recursive_mutex> lk(r.mu);
  if (every < 0) return CLARITY_INVALID_ARGUMENT;
  r.fault_every = every;
  r.alloc_count = 0;
  return CLARITY_SUCCESS;
}

}  // extern "C"
