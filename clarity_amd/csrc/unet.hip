// Learned warm start: a 3-level 3D U-Net (channels 16/32/64, 3x3x3
// convolutions, ReLU, skip connections) that maps an observation to an
// estimate of its 200-iteration Richardson-Lucy result; the estimate then
// seeds a short RL run.  No counterpart in the reference.
//
// Tensors are channels-last bfloat16 ([z][y][x][c], c fastest); weights are
// float32 masters with bfloat16 copies in two layouts, [tap][co][ci] for the
// forward pass and [tap][ci][co] for the input gradient.  Every 3x3x3
// convolution is an implicit GEMM (float32 accumulation) over K = 27 taps x
// input channels in chunks of 16: the forward pass stages input tiles in
// LDS and runs v_mfma_f32_16x16x32_bf16 on two taps at a time, the input
// gradient v_mfma_f32_16x16x16_bf16 on one.  The weight gradient is the same
// GEMM with K = voxels, staged through LDS, reduced over voxel chunks in a
// fixed order.
//
// Network (C = channels):
//   in(16: obs/mean, sigma_xy/4, sigma_z/8, zeros) -> c1a(16) -> e1(16)
//   pool -> c2a(32) -> e2(32) -> pool -> c3a(64) -> b3(64)
//   up(b3)|e2 (96) -> c4a(32) -> d2(32) -> up(d2)|e1 (48) -> c5a(16) -> d1(16)
//   out = obs/mean + head(d1) (1x1x1, 16 -> 1, float32)
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace clarity {
namespace {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

__device__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((unsigned)h << 16); }
// f2bf(a) | f2bf(b) << 16 in one instruction (v_cvt_pk_bf16_f32, round to
// nearest even: the same bits as f2bf for every non-NaN value; the
// activations it rounds are ReLU outputs and sums of them, never NaN)
typedef float f2v_t __attribute__((ext_vector_type(2)));
typedef __bf16 b2v_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector((f2v_t){a, b}, b2v_t));
}
// max(r, 0) with NaN -> +0 and -0 -> +0 (a stored activation is never -0)
__device__ __forceinline__ float relu(float r) { return r > 0.f ? r : 0.f; }

struct Vol {
  int D, H, W;
  __host__ __device__ long long n() const { return (long long)D * H * W; }
};

__device__ __forceinline__ void tap_offset(int tap, int &dz, int &dy, int &dx) {
  dz = tap / 9 - 1;
  dy = (tap / 3) % 3 - 1;
  dx = tap % 3 - 1;
}

// ------------------------------------------------------------------ conv

// Y = relu(conv(X, W) + b), one plane per block (the 64-output-channel
// layers; k_conv_ip and k_conv_co below take the others).  A block computes
// a 64 (x) by 4 (y) tile of output voxels of one z plane, one wavefront per
// row, all COUT channels.
// Input channels go in chunks of 16: the chunk's (3 x 6 x 66)-voxel input
// tile is staged in LDS as two planes of channel octets ([octet][voxel],
// 16 B each, one 16-byte read per MFMA operand), and every tap reads its
// fragments from there.  The MFMA computes the transposed product: operand
// A = weights (row = output channel, k = (tap, input channel)), operand B =
// input (column = voxel), D row = output channel 4*(lane/16) + j, column =
// voxel lane%16, so each lane stores four consecutive output channels of
// one voxel (8 B).
//
// With UP the input is the channel concatenation up(Xc) | S of a coarse
// tensor Xc (CA channels, nearest-neighbour 2x upsampled on the fly) and a
// fine tensor S (CIN - CA channels); otherwise X (CIN channels).
constexpr int kXT = 64, kYT = 4;                                // output tile
constexpr int kTX = kXT + 2, kTY = kYT + 2, kTZ = 3;             // input tile
constexpr int kTV = kTX * kTY * kTZ;                            // 1188 voxels

// The network's output layer fused into the last convolution (inference):
// out[v] = obs[v]/mean + b + sum_c w[c] d1[v][c], d1 rounded to bfloat16 as
// the stored activation would be; divided back into the observation's
// units (after a ReLU) when `rescale`.
struct HeadArgs {
  const float *obs = nullptr, *inv_mean = nullptr, *w = nullptr;  // w: 16 + bias
  float *out = nullptr;
  int rescale = 0;
};

// Epilogue of an output tile: y = relu(acc + bias) for the wavefront's row
// (D row = channel 4*(lane/16) + j, column = voxel lane%16 of each of the
// four 16-voxel groups), stored as bfloat16.
template <int COUT>
__device__ __forceinline__ void conv_epilogue(const f32x4 (&acc)[4][COUT / 16],
                                              const float *__restrict__ bias,
                                              bf16 *__restrict__ Y, Vol v, int x0, int y, int z,
                                              int m, int kq) {
  const long long rowbase = ((long long)z * v.H + y) * v.W;
#pragma unroll
  for (int ct = 0; ct < COUT / 16; ct++) {
    const int co = ct * 16 + 4 * kq;
    const float4 bb = *reinterpret_cast<const float4 *>(bias + co);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const int xo = x0 + mt * 16 + m;
      if (xo < v.W) {
        const float r0 = relu(acc[mt][ct][0] + bb.x), r1 = relu(acc[mt][ct][1] + bb.y);
        const float r2 = relu(acc[mt][ct][2] + bb.z), r3 = relu(acc[mt][ct][3] + bb.w);
        uint2 o;
        o.x = pk_bf16(r0, r1);
        o.y = pk_bf16(r2, r3);
        *reinterpret_cast<uint2 *>(Y + (rowbase + xo) * COUT + co) = o;
      }
    }
  }
}

template <int CIN, int COUT, int CA, bool UP>
__global__ void __launch_bounds__(256)
    k_conv_fwd(const bf16 *__restrict__ X, const bf16 *__restrict__ S,
               const bf16 *__restrict__ Wf, const float *__restrict__ bias,
               bf16 *__restrict__ Y, Vol v) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0 && (!UP || (CA % 16 == 0 && CA < CIN)),
                "channel chunks of 16");
  __shared__ uint4 tile[2][kTV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int x0 = blockIdx.x * kXT, y0 = blockIdx.y * kYT, z = blockIdx.z;
  f32x4 acc[4][COUT / 16];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int ct = 0; ct < COUT / 16; ct++) acc[mt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int c0 = 0; c0 < CIN; c0 += 16) {
    // stage: two 16-byte loads (8 channels each) per tile voxel.  All of a
    // thread's loads are issued before its first LDS store (the counter of
    // outstanding loads is in order, so a store loop that waits for each
    // load in turn serialises them); addresses are clamped into the volume
    // and values outside it cleared by a bit mask, not a branch around the
    // load.  With CA a multiple of 16 a chunk lies wholly in the upsampled
    // or the fine part of an UP input.
    constexpr int NI = (2 * kTV + 255) / 256;
    const bool coarse = UP && c0 < CA;
    uint4 stg[NI];
#pragma unroll
    for (int j = 0; j < NI; j++) {
      const int i0 = threadIdx.x + 256 * j;
      const int i = i0 < 2 * kTV ? i0 : 2 * kTV - 1;
      const int vox = i >> 1, half = i & 1;
      const int tx = vox % kTX, t2 = vox / kTX;
      const int ty = t2 % kTY, tz = t2 / kTY;
      const int gx = x0 + tx - 1, gy = y0 + ty - 1, gz = z + tz - 1;
      // bitwise, not short-circuit: no branches between the loads
      const bool in = (gx >= 0) & (gx < v.W) & (gy >= 0) & (gy < v.H) & (gz >= 0) & (gz < v.D);
      const int cx = min(max(gx, 0), v.W - 1), cy = min(max(gy, 0), v.H - 1);
      const int cz = min(max(gz, 0), v.D - 1);
      const int c = c0 + 8 * half;
      const bf16 *p;
      if (!UP) {
        p = X + (((long long)cz * v.H + cy) * v.W + cx) * CIN + c;
      } else if (coarse) {
        const int H2 = v.H >> 1, W2 = v.W >> 1;
        p = X + (((long long)(cz >> 1) * H2 + (cy >> 1)) * W2 + (cx >> 1)) * CA + c;
      } else {
        p = S + (((long long)cz * v.H + cy) * v.W + cx) * (CIN - CA) + (c - CA);
      }
      const uint4 val = *reinterpret_cast<const uint4 *>(p);
      const unsigned m = in ? 0xffffffffu : 0u;
      stg[j] = make_uint4(val.x & m, val.y & m, val.z & m, val.w & m);
    }
    __syncthreads();  // every wavefront is done with the previous chunk
#pragma unroll
    for (int j = 0; j < NI; j++) {
      const int i = threadIdx.x + 256 * j;
      if (i < 2 * kTV) tile[i & 1][i >> 1] = stg[j];
    }
    __syncthreads();
    // K = 32 per MFMA: two taps of the chunk's 16 channels.  Lane group
    // g = lane/16 supplies tap 2p + g/2, channels 8 (g % 2) .. + 7, in both
    // operands, so whatever order the hardware gives the 32 k, weights and
    // inputs pair up; the 28th tap has zero weights (its inputs are read at
    // tap 26's offsets).
    const int oct = kq & 1;
    // unrolled by two, not fully: a full unroll of the 14 pairs holds the
    // next chunk's loads and weights in ~50 more VGPRs (one wavefront per
    // SIMD less) for the layers with several input chunks
#pragma unroll 2
    for (int p = 0; p < 14; p++) {
      const int tp = 2 * p + (kq >> 1);
      const bool real = tp < 27;
      int dz, dy, dx;
      tap_offset(real ? tp : 26, dz, dy, dx);
      bf16x8 w[COUT / 16];
#pragma unroll
      for (int ct = 0; ct < COUT / 16; ct++) {
        const uint4 wv = *reinterpret_cast<const uint4 *>(
            Wf + ((long long)(real ? tp : 26) * COUT + ct * 16 + m) * CIN + c0 + 8 * oct);
        const unsigned z0 = real ? 0xffffffffu : 0u;
        w[ct] = __builtin_bit_cast(bf16x8, make_uint4(wv.x & z0, wv.y & z0, wv.z & z0, wv.w & z0));
      }
      const int rb = ((dz + 1) * kTY + wave + dy + 1) * kTX + dx + 1;
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const bf16x8 a = __builtin_bit_cast(bf16x8, tile[oct][rb + mt * 16 + m]);
#pragma unroll
        for (int ct = 0; ct < COUT / 16; ct++)
          acc[mt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[ct], a, acc[mt][ct], 0, 0, 0);
      }
    }
  }
  const int y = y0 + wave;
  if (y >= v.H) return;  // a whole wavefront (one row)
  conv_epilogue<COUT>(acc, bias, Y, v, x0, y, z, m, kq);
}

// Staging of one input plane (a 66 x 6 patch of a 16-channel chunk) for the
// plane-by-plane convolution below: two LDS buffers of [octet][voxel].
constexpr int kPV = kTX * kTY;               // 396 voxels per staged plane
constexpr int kPL = (2 * kPV + 255) / 256;   // 16-byte loads per thread and plane
// 4 voxels (64 B) between the two octet planes in LDS: unpadded, the
// 6336-byte plane stride sets the two octets' fragment reads on the same
// banks (0.31-0.35 conflict cycles per active LDS cycle, 0.12-0.14 padded;
// 8 voxels map like none, 1, 2 or 6 gain nothing)
constexpr int kPad = 4;
constexpr int kPS = kPV + kPad;              // LDS stride of an octet plane

// The thread's staging slots of a plane: byte offsets within a plane (x, y
// clamped into the volume; for an UP input also those into the coarse
// tensor's plane) and which slots lie inside the volume in x and y.  They
// are the same for every plane and chunk; a plane's base address is
// uniform, so each load is one scalar base plus one 32-bit offset.
struct ZmSlots {
  unsigned off[kPL], offc[kPL];
  unsigned inxy;  // bit j: slot j inside the volume in x and y
};

template <int CIN, int CA, bool UP>
__device__ __forceinline__ ZmSlots zm_slots(Vol v, int x0, int y0) {
  ZmSlots t;
  t.inxy = 0;
  constexpr int CF = UP ? CIN - CA : CIN;  // channels of the full-resolution tensor
#pragma unroll
  for (int j = 0; j < kPL; j++) {
    const int i0 = threadIdx.x + 256 * j;
    const int i = i0 < 2 * kPV ? i0 : 2 * kPV - 1;
    const int vox = i >> 1, half = i & 1;
    const int tx = vox % kTX, ty = vox / kTX;
    const int gx = x0 + tx - 1, gy = y0 + ty - 1;
    const bool in = (gx >= 0) & (gx < v.W) & (gy >= 0) & (gy < v.H);
    const int cx = min(max(gx, 0), v.W - 1), cy = min(max(gy, 0), v.H - 1);
    t.inxy |= (in ? 1u : 0u) << j;
    t.off[j] = 2u * ((unsigned)(cy * v.W + cx) * CF + 8 * half);
    t.offc[j] = UP ? 2u * ((unsigned)((cy >> 1) * (v.W >> 1) + (cx >> 1)) * CA + 8 * half) : 0u;
  }
  return t;
}

// the thread's share of input plane z = gz of chunk c0, masked to zero
// outside the volume
template <int CIN, int CA, bool UP>
__device__ __forceinline__ void zm_load(uint4 (&r)[kPL], const ZmSlots &t,
                                        const bf16 *__restrict__ X, const bf16 *__restrict__ S,
                                        Vol v, int gz, int c0) {
  const bool coarse = UP && c0 < CA;
  const bool zin = (gz >= 0) & (gz < v.D);
  const long long cz = min(max(gz, 0), v.D - 1);
  const long long plane = (long long)v.H * v.W;
  const char *pb;
  if (!UP)
    pb = reinterpret_cast<const char *>(X + cz * plane * CIN + c0);
  else if (coarse)
    pb = reinterpret_cast<const char *>(X + (cz >> 1) * (plane >> 2) * CA + c0);
  else
    pb = reinterpret_cast<const char *>(S + cz * plane * (CIN - CA) + (c0 - CA));
#pragma unroll
  for (int j = 0; j < kPL; j++) {
    const uint4 val = *reinterpret_cast<const uint4 *>(pb + (coarse ? t.offc[j] : t.off[j]));
    const unsigned m = (zin & ((t.inxy >> j) & 1u)) ? 0xffffffffu : 0u;
    r[j] = make_uint4(val.x & m, val.y & m, val.z & m, val.w & m);
  }
}

__device__ __forceinline__ void zm_store(const uint4 (&r)[kPL], uint4 *ring, int slot) {
#pragma unroll
  for (int j = 0; j < kPL; j++) {
    const int i = threadIdx.x + 256 * j;
    if (i < 2 * kPV) ring[(slot * 2 + (i & 1)) * kPS + (i >> 1)] = r[j];
  }
}

// Per-lane constants of an output tile's epilogue (16 output channels):
// the bias of the lane's four channels and, for the head, its weights, bias
// and 1/mean -- loaded once, before the main loop, so that no load is
// issued between the tile's stores (each would wait for the stores before it).
struct Epi16 {
  float4 bb, ww;
  float hb, im;
};

template <bool HEAD>
__device__ __forceinline__ Epi16 epi16_load(const float *__restrict__ bias, const HeadArgs &h,
                                            int kq) {
  Epi16 e;
  e.bb = *reinterpret_cast<const float4 *>(bias + 4 * kq);
  if constexpr (HEAD) {
    e.ww = *reinterpret_cast<const float4 *>(h.w + 4 * kq);
    e.hb = h.w[16];
    e.im = h.inv_mean[0];
  } else {
    e.ww = float4{0.f, 0.f, 0.f, 0.f};
    e.hb = e.im = 0.f;
  }
  return e;
}

// Epilogue of NP consecutive output planes (z0 ..) of one row: y =
// relu(acc + bias) stored as bfloat16 (16 channels), or with HEAD reduced
// over the channels by the head (each lane sums its 4 channels, two lane
// exchanges sum the 4 lanes of a voxel); the head's observation loads are
// issued before any store.
template <int NP, bool HEAD>
__device__ __forceinline__ void epi16_store(const f32x4 (&acc)[NP][4], const Epi16 &e,
                                            bf16 *__restrict__ Y, const HeadArgs &h, Vol v,
                                            int x0, int y, int z0, int m, int kq) {
  if constexpr (HEAD) {
    float ov[NP][4];
#pragma unroll
    for (int o = 0; o < NP; o++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const int xo = min(x0 + mt * 16 + m, v.W - 1), zo = min(z0 + o, v.D - 1);
        ov[o][mt] = h.obs[((long long)zo * v.H + y) * v.W + xo];
      }
    // each lane's 4-channel share of the head, then the two lane exchanges
    // for all NP x 4 values at once (one LDS round trip per stage, not 2
    // per value)
    float pp[NP][4];
#pragma unroll
    for (int o = 0; o < NP; o++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const unsigned q01 = pk_bf16(relu(acc[o][mt][0] + e.bb.x), relu(acc[o][mt][1] + e.bb.y));
        const unsigned q23 = pk_bf16(relu(acc[o][mt][2] + e.bb.z), relu(acc[o][mt][3] + e.bb.w));
        const float d0 = __uint_as_float(q01 << 16), d1 = __uint_as_float(q01 & 0xffff0000u);
        const float d2 = __uint_as_float(q23 << 16), d3 = __uint_as_float(q23 & 0xffff0000u);
        pp[o][mt] = e.ww.x * d0 + e.ww.y * d1 + e.ww.z * d2 + e.ww.w * d3;
      }
#pragma unroll
    for (int sh = 16; sh <= 32; sh *= 2) {
      float q[NP][4];
#pragma unroll
      for (int o = 0; o < NP; o++)
#pragma unroll
        for (int mt = 0; mt < 4; mt++) q[o][mt] = __shfl_xor(pp[o][mt], sh);
#pragma unroll
      for (int o = 0; o < NP; o++)
#pragma unroll
        for (int mt = 0; mt < 4; mt++) pp[o][mt] += q[o][mt];
    }
#pragma unroll
    for (int o = 0; o < NP; o++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const int xo = x0 + mt * 16 + m;
        if (kq == 0 && xo < v.W && z0 + o < v.D) {
          float r = ov[o][mt] * e.im + (e.hb + pp[o][mt]);
          if (h.rescale) r = (r > 0.f ? r : 0.f) / e.im;
          h.out[((long long)(z0 + o) * v.H + y) * v.W + xo] = r;
        }
      }
  } else {
#pragma unroll
    for (int o = 0; o < NP; o++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const int xo = x0 + mt * 16 + m;
        if (xo < v.W && z0 + o < v.D) {
          const float r0 = relu(acc[o][mt][0] + e.bb.x), r1 = relu(acc[o][mt][1] + e.bb.y);
          const float r2 = relu(acc[o][mt][2] + e.bb.z), r3 = relu(acc[o][mt][3] + e.bb.w);
          uint2 q;
          q.x = pk_bf16(r0, r1);
          q.y = pk_bf16(r2, r3);
          *reinterpret_cast<uint2 *>(Y + (((long long)(z0 + o) * v.H + y) * v.W + xo) * 16 +
                                     4 * kq) = q;
        }
      }
  }
}

template <int N, int I = 0, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<N, I + 1>(f);
  }
}

// Input-plane-major convolution for 16 output channels (the full-resolution
// layers): a block computes the 64 x 4 tile of kIB consecutive z planes
// and walks the kIB + 2 input planes it reads once each, chunk by chunk.
// An input plane feeds three output planes (dz = -1, 0, 1) through the same
// in-plane tap positions, so the K = 32 pairs are formed within one dz:
// pair p, lane half e = kq / 2 holds in-plane tap t = 2p + e (dy = t / 3 -
// 1, dx = t % 3 - 1; the tenth tap is zero), and each 16-voxel fragment
// read from LDS feeds the three planes' MFMAs (60 MFMAs per 20 fragment
// reads).  (plane, chunk) units go through two LDS buffers, the loads of
// the unit kIPDepth ahead in flight during the current unit's MFMAs; pair
// p + 1's LDS reads are issued before pair p's MFMAs.  Weights: one chunk
// in registers (60 VGPRs), more chunks in LDS ([chunk][dz][pair][lane], 15
// KB each).  Every output plane keeps its accumulators to the end and the
// epilogues follow the unit loop (a store inside it, behind the bounds
// branch, would make the wait for a unit's loads also wait for the loads
// issued after it: vmcnt counts loads and stores in order).  Per output
// voxel the sum runs over input planes z-1, z, z+1, then chunks, then
// pairs; where an upsampled chunk's two fine planes share a unit (below),
// their pairs interleave.
constexpr int kIB = 4;        // output planes per block
constexpr int kIPDepth = 2;   // units whose loads are in flight
constexpr bool kIPWLds = false;  // weights in LDS also for one chunk

// k_conv_ip's units: (input slot s, chunk c), slot-major; slot s is input
// plane z0 - 1 + s.  For an upsampled chunk the fine planes z0 + 2k and
// z0 + 2k + 1 (slots 2k + 1 and 2k + 2; z0 is a multiple of the even kIB)
// are one coarse plane, so one unit stages it once and feeds both slots'
// output planes (two); plain chunks keep one unit per slot.
struct IPUnit {
  int s, c;
  bool two;
};
template <int NCH, int NUP>
constexpr IPUnit ip_unit(int u) {
  int k = 0;
  for (int s = 0; s < kIB + 2; s++)
    for (int c = 0; c < NCH; c++) {
      const bool up = c < NUP;
      if (up && s % 2 == 0 && s >= 2 && s <= kIB) continue;  // merged into slot s - 1
      if (k == u) return IPUnit{s, c, up && s % 2 == 1 && s < kIB};
      k++;
    }
  return IPUnit{0, 0, false};
}

// POOL: also the 2x2x2 average pooling of the stored bfloat16 output into P
// at half resolution (each channel sums its eight inputs in the order
// k = 0..7 of the z, y, x offset bits, then scales by 1/8); needs
// v.H % 4 == 0, so every row of a tile is inside the volume.
template <int CIN, int CA, bool UP, bool HEAD, bool POOL = false>
__global__ void __launch_bounds__(256)
    k_conv_ip(const bf16 *__restrict__ X, const bf16 *__restrict__ S,
              const bf16 *__restrict__ Wf, const float *__restrict__ bias,
              bf16 *__restrict__ Y, Vol v, HeadArgs h, bf16 *__restrict__ P = nullptr) {
  static_assert(CIN % 16 == 0 && (!UP || (CA % 16 == 0 && CA < CIN)), "channel chunks of 16");
  static_assert(kIB % 2 == 0, "coarse planes pair up");
  constexpr int COUT = 16, NCH = CIN / 16, NUP = UP ? CA / 16 : 0;
  constexpr int NU = (kIB + 2) * NCH - NUP * (kIB / 2);
  constexpr bool WLDS = NCH > 1 || kIPWLds;
  __shared__ uint4 buf[2 * 2 * kPS];
  __shared__ uint4 wl[WLDS ? NCH * 15 * 64 : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int oct = kq & 1, e = kq >> 1;
  const int x0 = blockIdx.x * kXT, y0 = blockIdx.y * kYT, z0 = blockIdx.z * kIB;
  // weight fragment of chunk c, offset dz, pair p for this lane
  auto wfrag = [&](int c, int dz, int p) -> uint4 {
    const int t = 2 * p + e;
    const int tap = (dz + 1) * 9 + (t < 9 ? t : 8);
    const uint4 wv = *reinterpret_cast<const uint4 *>(
        Wf + ((long long)tap * COUT + m) * CIN + 16 * c + 8 * oct);
    const unsigned z = t < 9 ? 0xffffffffu : 0u;
    return make_uint4(wv.x & z, wv.y & z, wv.z & z, wv.w & z);
  };
  bf16x8 w[WLDS ? 1 : 15];
  if constexpr (WLDS) {
    for (int i = threadIdx.x; i < NCH * 15 * 64; i += 256) {
      const int ln = i & 63, f = i >> 6;  // f = (c * 3 + dz + 1) * 5 + p
      const int p = f % 5, dz = (f / 5) % 3 - 1, c = f / 15;
      const int mm = ln & 15, kk = ln >> 4;
      const int t = 2 * p + (kk >> 1);
      const int tap = (dz + 1) * 9 + (t < 9 ? t : 8);
      const uint4 wv = *reinterpret_cast<const uint4 *>(
          Wf + ((long long)tap * COUT + mm) * CIN + 16 * c + 8 * (kk & 1));
      const unsigned z = t < 9 ? 0xffffffffu : 0u;
      wl[i] = make_uint4(wv.x & z, wv.y & z, wv.z & z, wv.w & z);
    }
  } else {
#pragma unroll
    for (int f = 0; f < 15; f++) w[f] = __builtin_bit_cast(bf16x8, wfrag(0, f / 5 - 1, f % 5));
  }
  // the lane's fragment offset within a buffer for pair p (in-plane tap of
  // its lane half; the zero tenth tap reads tap 8's voxels)
  int foff[5];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    const int t0 = 2 * p, t1 = 2 * p + 1 < 9 ? 2 * p + 1 : 8;
    const int o0 = (t0 / 3) * kTX + t0 % 3, o1 = (t1 / 3) * kTX + t1 % 3;
    foff[p] = oct * kPS + wave * kTX + m + (e ? o1 : o0);
  }
  const ZmSlots t = zm_slots<CIN, CA, UP>(v, x0, y0);
  const int y = y0 + wave;
  const Epi16 ep = epi16_load<HEAD>(bias, h, kq);
  f32x4 acc[kIB][4];
  if constexpr (NUP > 0) {  // a merged unit starts two output planes
#pragma unroll
    for (int o = 0; o < kIB; o++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) acc[o][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  uint4 r[kIPDepth][kPL];
  static_for<kIPDepth>([&](auto uc) __attribute__((always_inline)) {
    constexpr int u = decltype(uc)::value;
    constexpr IPUnit U = ip_unit<NCH, NUP>(u);
    if constexpr (u < NU) zm_load<CIN, CA, UP>(r[u], t, X, S, v, z0 + U.s - 1, 16 * U.c);
  });
  // units one by one at compile time (a loop of 18 is not always unrolled,
  // and a unit index known only at run time would index register arrays)
  static_for<NU>([&](auto uc) __attribute__((always_inline)) {
    constexpr int u = decltype(uc)::value;
    constexpr IPUnit U = ip_unit<NCH, NUP>(u);
    constexpr int s = U.s, c = U.c, ns = U.two ? 2 : 1;  // slots s .. s + ns - 1, chunk c
    zm_store(r[u % kIPDepth], buf, u & 1);
    __syncthreads();  // the unit is in LDS; (WLDS) so are the weights
    if constexpr (u + kIPDepth < NU) {
      constexpr IPUnit Un = ip_unit<NCH, NUP>(u + kIPDepth);
      zm_load<CIN, CA, UP>(r[u % kIPDepth], t, X, S, v, z0 + Un.s - 1, 16 * Un.c);
    }
    if constexpr (NUP == 0 && c == 0 && s < kIB) {  // output plane s is first read here
#pragma unroll
      for (int mt = 0; mt < 4; mt++) acc[s][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    const uint4 *bb = buf + (u & 1) * 2 * kPS;
    // software pipeline over the five pairs: pair p + 1's LDS reads are
    // issued before pair p's MFMAs (scheduling barriers keep that order;
    // left alone, the compiler reads each fragment just before its MFMA
    // and waits out the LDS latency every time)
    bf16x8 fa[2][4], wa[2][3];
#define CLARITY_IP_READ_PAIR(P, B)                                                        \
  {                                                                                       \
    _Pragma("unroll") for (int mt = 0; mt < 4; mt++) fa[B][mt] =                          \
        __builtin_bit_cast(bf16x8, bb[foff[P] + mt * 16]);                                \
    _Pragma("unroll") for (int dz = -1; dz <= 1; dz++) {                                  \
      /* output plane z0 + o reads slot s + k at dz */                                    \
      const int olo = s - 1 - dz, ohi = olo + ns - 1;                                     \
      if (ohi < 0 || olo >= kIB) continue;                                                \
      if constexpr (WLDS)                                                                 \
        wa[B][dz + 1] = __builtin_bit_cast(bf16x8, wl[((c * 3 + dz + 1) * 5 + (P)) * 64 + lane]); \
    }                                                                                     \
  }
    CLARITY_IP_READ_PAIR(0, 0)
#pragma unroll
    for (int p = 0; p < 5; p++) {
      if (p + 1 < 5) CLARITY_IP_READ_PAIR(p + 1, (p + 1) & 1)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int dz = -1; dz <= 1; dz++)
#pragma unroll
        for (int k = 0; k < ns; k++) {
          const int o = s + k - 1 - dz;
          if (o < 0 || o >= kIB) continue;
#pragma unroll
          for (int mt = 0; mt < 4; mt++)
            acc[o][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                WLDS ? wa[p & 1][dz + 1] : w[(dz + 1) * 5 + p], fa[p & 1][mt], acc[o][mt], 0, 0,
                0);
        }
      __builtin_amdgcn_sched_barrier(0);
    }
#undef CLARITY_IP_READ_PAIR
  });
  // epilogues after the last unit: a store inside the unit loop, behind the
  // row/plane bounds branch, would make the wait for a unit's loads also
  // wait for the loads issued after it (vmcnt counts loads and stores in
  // order, and the count past a branch is taken as the smaller one)
  if constexpr (!POOL) {
    if (y >= v.H) return;
    epi16_store<kIB, HEAD>(acc, ep, Y, h, v, x0, y, z0, m, kq);
  } else {
    static_assert(!HEAD && kIB % 2 == 0, "pooling of the stored activation");
    epi16_store<kIB, false>(acc, ep, Y, h, v, x0, y, z0, m, kq);
    // the tile's bfloat16 outputs, two planes at a time, staged in LDS as
    // [plane][row][x parity][x / 2][channel quad] (8 B each; the unit
    // buffers are free now), then one pooled voxel's channel quad per thread
    uint2 *st = reinterpret_cast<uint2 *>(buf);
    static_assert(2 * kYT * kXT * 4 * 8 <= (int)sizeof(buf), "staging fits");
    const int W2 = v.W >> 1, H2 = v.H >> 1;
    const int quad = threadIdx.x & 3, vox = threadIdx.x >> 2;  // 64 pooled voxels
    const int px = vox & 31, py = vox >> 5;
#pragma unroll
    for (int pp = 0; pp < kIB / 2; pp++) {
      __syncthreads();  // the buffers' last readers are done
#pragma unroll
      for (int oo = 0; oo < 2; oo++)
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
          const f32x4 &a = acc[2 * pp + oo][mt];
          uint2 q;
          q.x = pk_bf16(relu(a[0] + ep.bb.x), relu(a[1] + ep.bb.y));
          q.y = pk_bf16(relu(a[2] + ep.bb.z), relu(a[3] + ep.bb.w));
          // even x in the first half of a row, odd x in the second: the
          // pooling's reads of consecutive pooled voxels are then contiguous
          const int xs = mt * 16 + m;
          st[((oo * kYT + wave) * kXT + (xs & 1) * (kXT / 2) + (xs >> 1)) * 4 + kq] = q;
        }
      __syncthreads();
      float sum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint2 q = st[(((k >> 2) * kYT + 2 * py + ((k >> 1) & 1)) * kXT + (k & 1) * (kXT / 2) + px) * 4 + quad];
        sum[0] += bf2f((bf16)(q.x & 0xffffu));
        sum[1] += bf2f((bf16)(q.x >> 16));
        sum[2] += bf2f((bf16)(q.y & 0xffffu));
        sum[3] += bf2f((bf16)(q.y >> 16));
      }
      const int gx = (x0 >> 1) + px, gy = (y0 >> 1) + py, gz = (z0 >> 1) + pp;
      if (gx < W2 && 2 * gz < v.D) {
        uint2 o;
        o.x = pk_bf16(0.125f * sum[0], 0.125f * sum[1]);
        o.y = pk_bf16(0.125f * sum[2], 0.125f * sum[3]);
        *reinterpret_cast<uint2 *>(P + (((long long)gz * H2 + gy) * W2 + gx) * 16 + 4 * quad) = o;
      }
    }
  }
}

constexpr int kCoDepth = 1;  // units whose loads are in flight (k_conv_co)

// k_conv_co's units: (chunk c, input slot s), chunk-major; the slots of an
// upsampled chunk pair up as in k_conv_ip (slots 2k + 1 and 2k + 2 are one
// coarse plane: z0 is a multiple of the even IB).
template <int NCH, int NUP, int IB>
constexpr IPUnit co_unit(int u) {
  int k = 0;
  for (int c = 0; c < NCH; c++)
    for (int s = 0; s < IB + 2; s++) {
      const bool up = c < NUP;
      if (up && s % 2 == 0 && s >= 2 && s <= IB) continue;  // merged into slot s - 1
      if (k == u) return IPUnit{s, c, up && s % 2 == 1 && s < IB};
      k++;
    }
  return IPUnit{0, 0, false};
}

// The same plane-by-plane scheme for more output channels (COUT/16 tiles,
// each fragment feeding up to 3 x COUT/16 MFMAs), chunk-outer: the chunk's
// weights (COUT/16 x 15 KB) sit in LDS while its IB + 2 input planes pass,
// and every output plane of the block keeps its accumulators to the end.
// Per output voxel the sum runs over chunks, then input planes z-1, z, z+1,
// then pairs; an upsampled chunk's two fine planes that share a unit
// (co_unit) interleave by pair.
// POOL (IB = 2): also the 2x2x2 average pooling of the output into P, as
// k_conv_ip's (the level's sizes are even, so a pooled voxel's rows and
// planes are all inside the volume or all outside).
template <int CIN, int COUT, int CA, bool UP, int IB, bool POOL = false>
__global__ void __launch_bounds__(256)
    k_conv_co(const bf16 *__restrict__ X, const bf16 *__restrict__ S,
              const bf16 *__restrict__ Wf, const float *__restrict__ bias,
              bf16 *__restrict__ Y, Vol v, bf16 *__restrict__ P = nullptr) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0 && (!UP || (CA % 16 == 0 && CA < CIN)),
                "channel chunks of 16");
  static_assert(!UP || IB % 2 == 0, "coarse planes pair up");
  constexpr int NCT = COUT / 16, NCH = CIN / 16, NS = IB + 2;
  constexpr int NUP = UP ? CA / 16 : 0, NU = NCH * NS - NUP * (IB / 2);
  constexpr int NW = NCT * 15 * 64;  // weight fragments of a chunk
  __shared__ uint4 buf[2 * 2 * kPS];
  __shared__ uint4 wl[NW];  // [dz][pair][ct][lane]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int oct = kq & 1, e = kq >> 1;
  const int x0 = blockIdx.x * kXT, y0 = blockIdx.y * kYT, z0 = blockIdx.z * IB;
  int foff[5];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    const int t0 = 2 * p, t1 = 2 * p + 1 < 9 ? 2 * p + 1 : 8;
    const int o0 = (t0 / 3) * kTX + t0 % 3, o1 = (t1 / 3) * kTX + t1 % 3;
    foff[p] = oct * kPS + wave * kTX + m + (e ? o1 : o0);
  }
  const ZmSlots t = zm_slots<CIN, CA, UP>(v, x0, y0);
  const int y = y0 + wave;
  float4 bb[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ct++) bb[ct] = *reinterpret_cast<const float4 *>(bias + ct * 16 + 4 * kq);
  f32x4 acc[IB][NCT][4];
#pragma unroll
  for (int o = 0; o < IB; o++)
#pragma unroll
    for (int ct = 0; ct < NCT; ct++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) acc[o][ct][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
  uint4 r[kCoDepth][kPL];
  static_for<kCoDepth>([&](auto uc) __attribute__((always_inline)) {
    constexpr int u = decltype(uc)::value;
    constexpr IPUnit U = co_unit<NCH, NUP, IB>(u);
    if constexpr (u < NU) zm_load<CIN, CA, UP>(r[u], t, X, S, v, z0 - 1 + U.s, 16 * U.c);
  });
  static_for<NU>([&](auto uc) __attribute__((always_inline)) {
    constexpr int u = decltype(uc)::value;
    constexpr IPUnit U = co_unit<NCH, NUP, IB>(u);
    // chunk c, input slots s .. s + ns - 1 (input plane z0 - 1 + s)
    constexpr int c = U.c, s = U.s, ns = U.two ? 2 : 1;
    zm_store(r[u % kCoDepth], buf, u & 1);
    __syncthreads();  // the unit is in LDS; the previous chunk's MFMAs are done
    if constexpr (s == 0) {
      for (int i = threadIdx.x; i < NW; i += 256) {
        const int ln = i & 63, f = i >> 6;  // f = ((dz + 1) * 5 + p) * NCT + ct
        const int ct = f % NCT, p = (f / NCT) % 5, dz = f / (NCT * 5) - 1;
        const int mm = ln & 15, kk = ln >> 4;
        const int tt = 2 * p + (kk >> 1);
        const int tap = (dz + 1) * 9 + (tt < 9 ? tt : 8);
        const uint4 wv = *reinterpret_cast<const uint4 *>(
            Wf + ((long long)tap * COUT + ct * 16 + mm) * CIN + 16 * c + 8 * (kk & 1));
        const unsigned z = tt < 9 ? 0xffffffffu : 0u;
        wl[i] = make_uint4(wv.x & z, wv.y & z, wv.z & z, wv.w & z);
      }
      __syncthreads();
    }
    if constexpr (u + kCoDepth < NU) {
      constexpr IPUnit Un = co_unit<NCH, NUP, IB>(u + kCoDepth);
      zm_load<CIN, CA, UP>(r[u % kCoDepth], t, X, S, v, z0 - 1 + Un.s, 16 * Un.c);
    }
    const uint4 *bq = buf + (u & 1) * 2 * kPS;
    bf16x8 fa[2][4], wa[2][3][NCT];
#define CLARITY_CO_READ_PAIR(P, B)                                                        \
  {                                                                                       \
    _Pragma("unroll") for (int mt = 0; mt < 4; mt++) fa[B][mt] =                          \
        __builtin_bit_cast(bf16x8, bq[foff[P] + mt * 16]);                                \
    _Pragma("unroll") for (int dz = -1; dz <= 1; dz++) {                                  \
      const int olo = s - 1 - dz, ohi = olo + ns - 1;                                     \
      if (ohi < 0 || olo >= IB) continue;                                                 \
      _Pragma("unroll") for (int ct = 0; ct < NCT; ct++) wa[B][dz + 1][ct] =              \
          __builtin_bit_cast(bf16x8, wl[(((dz + 1) * 5 + (P)) * NCT + ct) * 64 + lane]);  \
    }                                                                                     \
  }
    CLARITY_CO_READ_PAIR(0, 0)
#pragma unroll
    for (int p = 0; p < 5; p++) {
      if (p + 1 < 5) CLARITY_CO_READ_PAIR(p + 1, (p + 1) & 1)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int dz = -1; dz <= 1; dz++)
#pragma unroll
        for (int k = 0; k < ns; k++) {
          const int o = s + k - 1 - dz;  // output plane z0 + o reads slot s + k at dz
          if (o < 0 || o >= IB) continue;
#pragma unroll
          for (int ct = 0; ct < NCT; ct++)
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
              acc[o][ct][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                  wa[p & 1][dz + 1][ct], fa[p & 1][mt], acc[o][ct][mt], 0, 0, 0);
        }
      __builtin_amdgcn_sched_barrier(0);
    }
#undef CLARITY_CO_READ_PAIR
  });
  if (!POOL && y >= v.H) return;
#pragma unroll
  for (int o = 0; o < IB; o++)
#pragma unroll
    for (int ct = 0; ct < NCT; ct++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const int xo = x0 + mt * 16 + m;
        if (xo < v.W && z0 + o < v.D && y < v.H) {
          const float r0 = relu(acc[o][ct][mt][0] + bb[ct].x), r1 = relu(acc[o][ct][mt][1] + bb[ct].y);
          const float r2 = relu(acc[o][ct][mt][2] + bb[ct].z), r3 = relu(acc[o][ct][mt][3] + bb[ct].w);
          uint2 q;
          q.x = pk_bf16(r0, r1);
          q.y = pk_bf16(r2, r3);
          *reinterpret_cast<uint2 *>(Y + (((long long)(z0 + o) * v.H + y) * v.W + xo) * COUT +
                                     ct * 16 + 4 * kq) = q;
        }
      }
  if constexpr (POOL) {
    static_assert(IB == 2, "one pooled plane per block");
    // per 16-channel tile: the two planes' bfloat16 outputs staged in LDS as
    // [plane][row][x parity][x / 2][channel quad], then one pooled voxel's
    // channel quad per thread, summed in k = 0..7 order
    uint2 *st = reinterpret_cast<uint2 *>(buf);
    static_assert(2 * kYT * kXT * 4 * 8 <= (int)sizeof(buf), "staging fits");
    const int W2 = v.W >> 1, H2 = v.H >> 1;
    const int quad = threadIdx.x & 3, vox = threadIdx.x >> 2;
    const int px = vox & 31, py = vox >> 5;
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      __syncthreads();  // the buffers' last readers are done
#pragma unroll
      for (int o = 0; o < 2; o++)
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
          const f32x4 &a = acc[o][ct][mt];
          uint2 q;
          q.x = pk_bf16(relu(a[0] + bb[ct].x), relu(a[1] + bb[ct].y));
          q.y = pk_bf16(relu(a[2] + bb[ct].z), relu(a[3] + bb[ct].w));
          const int xs = mt * 16 + m;
          st[((o * kYT + wave) * kXT + (xs & 1) * (kXT / 2) + (xs >> 1)) * 4 + kq] = q;
        }
      __syncthreads();
      float sum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint2 q = st[(((k >> 2) * kYT + 2 * py + ((k >> 1) & 1)) * kXT + (k & 1) * (kXT / 2) + px) * 4 + quad];
        sum[0] += bf2f((bf16)(q.x & 0xffffu));
        sum[1] += bf2f((bf16)(q.x >> 16));
        sum[2] += bf2f((bf16)(q.y & 0xffffu));
        sum[3] += bf2f((bf16)(q.y >> 16));
      }
      const int gx = (x0 >> 1) + px, gy = (y0 >> 1) + py, gz = z0 >> 1;
      if (gx < W2 && gy < H2 && 2 * gz < v.D) {
        uint2 o2;
        o2.x = pk_bf16(0.125f * sum[0], 0.125f * sum[1]);
        o2.y = pk_bf16(0.125f * sum[2], 0.125f * sum[3]);
        *reinterpret_cast<uint2 *>(P + (((long long)gz * H2 + gy) * W2 + gx) * COUT + ct * 16 +
                                   4 * quad) = o2;
      }
    }
  }
}

// The first convolution for inference: the network input (obs/mean,
// sigma_xy/4, sigma_z/8, then 13 zero channels) is built in the staging
// from the observation itself, and only its 4 leading channels are kept
// (8 B per tile voxel).  K = 32 per MFMA = 8 taps x 4 channels: lane group
// g supplies taps 8p + 2g and 8p + 2g + 1, so 4 MFMAs cover the 27 taps
// (5 with zero weights) where the 16-channel layout needs 14.  The same
// products as k_conv_fwd on the materialised input (its other channels are
// zero).  Weights Wf as k_conv_fwd's with CIN = 16 (16 output channels).
// A block computes the 64 x 4 tile of IB consecutive z planes from its
// IB + 2 input planes, all staged at once (IB = 8: 32 KB), so each input
// voxel is converted (IB + 2) / IB times instead of three times.
template <int IB>
__global__ void __launch_bounds__(256)
    k_conv_in(const float *__restrict__ obs, const float *__restrict__ inv_mean, float sxy,
              float sz, const bf16 *__restrict__ Wf, const float *__restrict__ bias,
              bf16 *__restrict__ Y, Vol v) {
  constexpr int COUT = 16;
  constexpr int NV = (IB + 2) * kPV;
  __shared__ uint2 tile[NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int x0 = blockIdx.x * kXT, y0 = blockIdx.y * kYT, z0 = blockIdx.z * IB;
  const float im = inv_mean[0];
  // channel 1 (sigma_xy/4) is packed with channel 0 below
  const unsigned w1 = (unsigned)f2bf(sz * 0.125f);  // channel 2 (3: zero)
  const float4 bb = *reinterpret_cast<const float4 *>(bias + 4 * kq);
  // weights of the four K = 32 groups (taps 8p + 2g, 8p + 2g + 1)
  bf16x8 w[4];
#pragma unroll
  for (int p = 0; p < 4; p++) {
    const int ta = 8 * p + 2 * kq, tb = ta + 1;
    const unsigned ma = ta < 27 ? 0xffffffffu : 0u, mb = tb < 27 ? 0xffffffffu : 0u;
    const int ca = ta < 27 ? ta : 26, cb = tb < 27 ? tb : 26;
    const uint2 wa = *reinterpret_cast<const uint2 *>(Wf + ((long long)ca * COUT + m) * 16);
    const uint2 wb = *reinterpret_cast<const uint2 *>(Wf + ((long long)cb * COUT + m) * 16);
    w[p] = __builtin_bit_cast(bf16x8, make_uint4(wa.x & ma, wa.y & ma, wb.x & mb, wb.y & mb));
  }
  constexpr int NI = (NV + 255) / 256;
  float ov[NI];
  bool inb[NI];
#pragma unroll
  for (int j = 0; j < NI; j++) {  // all loads first (clamped, no branches)
    const int i0 = threadIdx.x + 256 * j;
    const int i = i0 < NV ? i0 : NV - 1;
    const int tz = i / kPV, r = i - tz * kPV;
    const int ty = r / kTX, tx = r - ty * kTX;
    const int gx = x0 + tx - 1, gy = y0 + ty - 1, gz = z0 + tz - 1;
    inb[j] = (gx >= 0) & (gx < v.W) & (gy >= 0) & (gy < v.H) & (gz >= 0) & (gz < v.D);
    const int cx = min(max(gx, 0), v.W - 1), cy = min(max(gy, 0), v.H - 1);
    const int cz = min(max(gz, 0), v.D - 1);
    ov[j] = obs[((long long)cz * v.H + cy) * v.W + cx];
  }
#pragma unroll
  for (int j = 0; j < NI; j++) {
    const int i = threadIdx.x + 256 * j;
    const unsigned mk = inb[j] ? 0xffffffffu : 0u;
    if (i < NV) tile[i] = make_uint2(pk_bf16(ov[j] * im, sxy * 0.25f) & mk, w1 & mk);
  }
  __syncthreads();
  const int y = y0 + wave;
  if (y >= v.H) return;  // a whole wavefront (one row); no barrier follows
#pragma unroll
  for (int o = 0; o < IB; o++) {
    f32x4 acc[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int p = 0; p < 4; p++) {
      const int ta = 8 * p + 2 * kq, tb = ta + 1;
      const int ca = ta < 27 ? ta : 26, cb = tb < 27 ? tb : 26;
      int dz, dy, dx;
      tap_offset(ca, dz, dy, dx);
      const int ra = ((o + dz + 1) * kTY + wave + dy + 1) * kTX + dx + 1;
      tap_offset(cb, dz, dy, dx);
      const int rb = ((o + dz + 1) * kTY + wave + dy + 1) * kTX + dx + 1;
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const uint2 a0 = tile[ra + mt * 16 + m], a1 = tile[rb + mt * 16 + m];
        const bf16x8 a = __builtin_bit_cast(bf16x8, make_uint4(a0.x, a0.y, a1.x, a1.y));
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[p], a, acc[mt], 0, 0, 0);
      }
    }
    if (z0 + o >= v.D) break;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const int xo = x0 + mt * 16 + m;
      if (xo < v.W) {
        const float r0 = relu(acc[mt][0] + bb.x), r1 = relu(acc[mt][1] + bb.y);
        const float r2 = relu(acc[mt][2] + bb.z), r3 = relu(acc[mt][3] + bb.w);
        uint2 q;
        q.x = pk_bf16(r0, r1);
        q.y = pk_bf16(r2, r3);
        *reinterpret_cast<uint2 *>(Y + (((long long)(z0 + o) * v.H + y) * v.W + xo) * COUT +
                                   4 * kq) = q;
      }
    }
  }
}

// GX = conv^T(G * [Y > 0], W): input gradient of a ReLU convolution.
// Same GEMM shape with taps mirrored and W laid out [tap][ci][co].
template <int CIN, int COUT>
__global__ void __launch_bounds__(256)
    k_conv_bwd_data(const bf16 *__restrict__ G, const bf16 *__restrict__ Yo,
                    const bf16 *__restrict__ Wb, bf16 *__restrict__ GX, Vol v) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int x0 = blockIdx.x * 64 + wave * 16;
  const int y = blockIdx.y, z = blockIdx.z;
  const int xv = x0 + m;
  f32x4 acc[CIN / 16];
#pragma unroll
  for (int c = 0; c < CIN / 16; c++) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 27; tap++) {
    int dz, dy, dx;
    tap_offset(tap, dz, dy, dx);
    // output voxel p = q - off reads input q: GX[q] += W[tap] . G[q - off]
    const int zz = z - dz, yy = y - dy, xx = xv - dx;
    const bool in = xv < v.W && zz >= 0 && zz < v.D && yy >= 0 && yy < v.H && xx >= 0 &&
                    xx < v.W;
    const long long off = (((long long)zz * v.H + yy) * v.W + xx) * COUT + 4 * kq;
    const bf16 *wp = Wb + ((long long)tap * CIN + m) * COUT + 4 * kq;
#pragma unroll
    for (int c0 = 0; c0 < COUT; c0 += 16) {
      s16x4 a = {0, 0, 0, 0};
      if (in) {
        const s16x4 g = *reinterpret_cast<const s16x4 *>(G + off + c0);
        const s16x4 o = *reinterpret_cast<const s16x4 *>(Yo + off + c0);
#pragma unroll
        for (int j = 0; j < 4; j++) a[j] = (o[j] > 0) ? g[j] : (short)0;  // bf16 > 0
      }
#pragma unroll
      for (int it = 0; it < CIN / 16; it++) {
        const s16x4 b = *reinterpret_cast<const s16x4 *>(wp + it * 16 * COUT + c0);
        acc[it] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[it], 0, 0, 0);
      }
    }
  }
  const long long rowbase = ((long long)z * v.H + y) * v.W;
#pragma unroll
  for (int it = 0; it < CIN / 16; it++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int xo = x0 + 4 * kq + j;
      if (xo < v.W) GX[(rowbase + xo) * CIN + it * 16 + m] = f2bf(acc[it][j]);
    }
  }
}

// Weight gradient of one tap over one chunk of voxels:
// P[chunk][tap][co][ci] = sum_v (G * [Y > 0])[v][co] * X[v + off][ci].
// 64 voxels per step are staged in LDS (masked G and shifted X); wavefront
// w takes voxels 16w..16w+15 of each step as the K of its MFMAs; the four
// wavefronts' sums are added at the end.
constexpr int kWChunk = 8192;

template <int CIN, int COUT>
__global__ void __launch_bounds__(256)
    k_conv_bwd_w(const bf16 *__restrict__ G, const bf16 *__restrict__ Yo,
                 const bf16 *__restrict__ X, float *__restrict__ P, Vol v) {
  constexpr int NT = (COUT / 16) * (CIN / 16);
  __shared__ bf16 sg[64 * COUT];
  __shared__ bf16 sx[64 * CIN];
  __shared__ float red[4][16 * 16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int tap = blockIdx.y;
  int dz, dy, dx;
  tap_offset(tap, dz, dy, dx);
  const long long vb = (long long)blockIdx.x * kWChunk;
  const long long ve = vb + kWChunk < v.n() ? vb + kWChunk : v.n();
  f32x4 acc[NT];
#pragma unroll
  for (int i = 0; i < NT; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long long s0 = vb; s0 < ve; s0 += 64) {
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * COUT; i += 256) {
      const long long p = s0 + i / COUT;
      bf16 gv = 0;
      if (p < ve) {
        const long long o = p * COUT + i % COUT;
        if ((short)Yo[o] > 0) gv = G[o];
      }
      sg[i] = gv;
    }
    for (int i = threadIdx.x; i < 64 * CIN; i += 256) {
      const long long p = s0 + i / CIN;
      bf16 xv = 0;
      if (p < ve) {
        const int x = (int)(p % v.W);
        const long long t = p / v.W;
        const int y = (int)(t % v.H), z = (int)(t / v.H);
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        if (zz >= 0 && zz < v.D && yy >= 0 && yy < v.H && xx >= 0 && xx < v.W)
          xv = X[(((long long)zz * v.H + yy) * v.W + xx) * CIN + i % CIN];
      }
      sx[i] = xv;
    }
    __syncthreads();
    const int vr = wave * 16 + 4 * kq;  // the lane's 4 voxels
#pragma unroll
    for (int ct = 0; ct < COUT / 16; ct++) {
      s16x4 a;
#pragma unroll
      for (int j = 0; j < 4; j++) a[j] = (short)sg[(vr + j) * COUT + ct * 16 + n];
#pragma unroll
      for (int it = 0; it < CIN / 16; it++) {
        s16x4 b;
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = (short)sx[(vr + j) * CIN + it * 16 + n];
        acc[ct * (CIN / 16) + it] =
            __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[ct * (CIN / 16) + it], 0, 0, 0);
      }
    }
  }
  // D fragment: row (co) 4*kq + j, column (ci) n
  float *out = P + ((long long)blockIdx.x * 27 + tap) * COUT * CIN;
#pragma unroll
  for (int i = 0; i < NT; i++) {
    const int ct = i / (CIN / 16), it = i % (CIN / 16);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; j++) red[wave][(4 * kq + j) * 16 + n] = acc[i][j];
    __syncthreads();
    if (threadIdx.x < 256) {
      const int r = threadIdx.x / 16, c = threadIdx.x % 16;
      const float sum = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                        red[3][threadIdx.x];
      out[(ct * 16 + r) * CIN + it * 16 + c] = sum;
    }
  }
}

// Bias gradient partials: Pb[chunk][co] = sum_v (G * [Y > 0])[v][co].
template <int COUT>
__global__ void __launch_bounds__(256)
    k_bias_grad(const bf16 *__restrict__ G, const bf16 *__restrict__ Yo,
                float *__restrict__ Pb, long long nvox) {
  __shared__ float sh[256];
  const long long vb = (long long)blockIdx.x * kWChunk;
  const long long ve = vb + kWChunk < nvox ? vb + kWChunk : nvox;
  const int co = threadIdx.x % COUT;
  const int lanes = 256 / COUT;
  float acc = 0.f;
  for (long long p = vb + threadIdx.x / COUT; p < ve; p += lanes) {
    const long long o = p * COUT + co;
    if ((short)Yo[o] > 0) acc += bf2f(G[o]);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < COUT) {
    float s = 0.f;
    for (int k = 0; k < lanes; k++) s += sh[k * COUT + threadIdx.x];
    Pb[(long long)blockIdx.x * COUT + threadIdx.x] = s;
  }
}

// out[i] = sum_c P[c][i] for c < nchunks (fixed order)
__global__ void k_reduce_chunks(const float *__restrict__ P, int nchunks, long long n,
                                float *__restrict__ out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < nchunks; c++) s += P[(long long)c * n + i];
    out[i] = s;
  }
}

// ----------------------------------------------------- pool / up / concat

#define VOX_LOOP(i, n)                                                      \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (long long)gridDim.x * blockDim.x)

// GX (fine) = GX + pool^T(GP) = GX + GP[parent] / 8
__global__ void k_pool_bwd_add(const bf16 *__restrict__ GP, bf16 *__restrict__ GX, Vol v,
                               int C) {
  const long long n = v.n() * C;
  VOX_LOOP(i, n) {
    const int c = (int)(i % C);
    long long p = i / C;
    const int x = (int)(p % v.W);
    p /= v.W;
    const int y = (int)(p % v.H), z = (int)(p / v.H);
    const long long q = (((long long)(z / 2) * (v.H / 2) + y / 2) * (v.W / 2) + x / 2) * C + c;
    GX[i] = f2bf(bf2f(GX[i]) + 0.125f * bf2f(GP[q]));
  }
}

// Y (fine, Ctot channels) channels [0, C1) = nearest upsampling of X
// (coarse, C1 channels); channels [C1, Ctot) = S (fine, Ctot - C1 channels)
__global__ void k_up_concat(const bf16 *__restrict__ X, const bf16 *__restrict__ S,
                            bf16 *__restrict__ Y, Vol v, int C1, int Ctot) {
  const int CS = Ctot - C1;
  const long long n = v.n() * Ctot;
  VOX_LOOP(i, n) {
    const int c = (int)(i % Ctot);
    long long p = i / Ctot;
    const long long pv = p;
    const int x = (int)(p % v.W);
    p /= v.W;
    const int y = (int)(p % v.H), z = (int)(p / v.H);
    if (c < C1) {
      Y[i] = X[(((long long)(z / 2) * (v.H / 2) + y / 2) * (v.W / 2) + x / 2) * C1 + c];
    } else {
      Y[i] = S[pv * CS + (c - C1)];
    }
  }
}

// GXc (coarse, C1) = sum over the 8 children of G channels [0, C1);
// GS (fine, Ctot - C1) = G channels [C1, Ctot)
__global__ void k_up_concat_bwd(const bf16 *__restrict__ G, bf16 *__restrict__ GXc,
                                bf16 *__restrict__ GS, Vol v, int C1, int Ctot) {
  const int CS = Ctot - C1;
  const int D2 = v.D / 2, H2 = v.H / 2, W2 = v.W / 2;
  const long long nc = (long long)D2 * H2 * W2 * C1;
  VOX_LOOP(i, nc) {
    const int c = (int)(i % C1);
    long long p = i / C1;
    const int x = (int)(p % W2);
    p /= W2;
    const int y = (int)(p % H2), z = (int)(p / H2);
    float s = 0.f;
    for (int k = 0; k < 8; k++) {
      const int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
      s += bf2f(G[(((long long)zz * v.H + yy) * v.W + xx) * Ctot + c]);
    }
    GXc[i] = f2bf(s);
  }
  const long long nf = v.n() * CS;
  VOX_LOOP(i, nf) {
    const int c = (int)(i % CS);
    const long long p = i / CS;
    GS[i] = G[p * Ctot + C1 + c];
  }
}

// ------------------------------------------------------ input, head, loss

// in[v] = (obs/mean, sigma_xy/4, sigma_z/8, 0...) in 16 channels
__global__ void k_input(const float *__restrict__ obs, const float *__restrict__ inv_mean,
                        float sxy, float sz, bf16 *__restrict__ in, long long n) {
  const float im = inv_mean[0];
  VOX_LOOP(i, n) {
    bf16 *p = in + i * 16;
    p[0] = f2bf(obs[i] * im);
    p[1] = f2bf(sxy * 0.25f);
    p[2] = f2bf(sz * 0.125f);
    for (int c = 3; c < 16; c++) p[c] = 0;
  }
}

// out[v] = obs[v]/mean + b + sum_c w[c] d1[v][c]   (scaled by mean when
// `rescale`, i.e. in the observation's units)
__global__ void k_head(const bf16 *__restrict__ d1, const float *__restrict__ w,
                       const float *__restrict__ obs, const float *__restrict__ inv_mean,
                       float *__restrict__ out, long long n, int rescale) {
  const float im = inv_mean[0];
  const float b = w[16];
  VOX_LOOP(i, n) {
    float s = b;
    for (int c = 0; c < 16; c++) s += w[c] * bf2f(d1[i * 16 + c]);
    float r = obs[i] * im + s;
    if (rescale) r = (r > 0.f ? r : 0.f) / im;
    out[i] = r;
  }
}

// Loss L = mean((out - t/mean)^2) (fp64 partials); gout = 2 (out - t/mean)/n
// and the head's gradients: gd1[v][c] = gout w[c], partial dw, db.
__global__ void __launch_bounds__(256)
    k_loss_head_bwd(const float *__restrict__ out, const float *__restrict__ tgt,
                    const float *__restrict__ inv_mean, const bf16 *__restrict__ d1,
                    const float *__restrict__ w, bf16 *__restrict__ gd1,
                    double *__restrict__ lossp, float *__restrict__ whp, long long n) {
  __shared__ double sl[256];
  __shared__ float sw[17][256];
  const float im = inv_mean[0];
  const float inv_n = 1.0f / (float)n;
  double lacc = 0.0;
  float wacc[17];
  for (int c = 0; c < 17; c++) wacc[c] = 0.f;
  VOX_LOOP(i, n) {
    const float d = out[i] - tgt[i] * im;
    lacc += (double)d * d;
    const float g = 2.f * d * inv_n;
    for (int c = 0; c < 16; c++) {
      const float a = bf2f(d1[i * 16 + c]);
      wacc[c] += g * a;
      gd1[i * 16 + c] = f2bf(g * w[c]);
    }
    wacc[16] += g;
  }
  sl[threadIdx.x] = lacc;
  for (int c = 0; c < 17; c++) sw[c][threadIdx.x] = wacc[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sl[threadIdx.x] += sl[threadIdx.x + s];
      for (int c = 0; c < 17; c++) sw[c][threadIdx.x] += sw[c][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    lossp[blockIdx.x] = sl[0];
    for (int c = 0; c < 17; c++) whp[blockIdx.x * 17 + c] = sw[c][0];
  }
}

// Adam on float32 masters; refreshes the bf16 copies of conv weights in
// both layouts ([tap][co][ci] and [tap][ci][co]) when cin > 0.
__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ s,
                       const float *__restrict__ g, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, bf16 *__restrict__ wf,
                       bf16 *__restrict__ wb, int cin, int cout) {
  VOX_LOOP(i, n) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float si = b2 * s[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    s[i] = si;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(si / c2) + eps);
    p[i] = pi;
    if (cin > 0) {
      // i indexes [tap][co][ci]
      const int ci = (int)(i % cin);
      const long long t = i / cin;
      const int co = (int)(t % cout), tap = (int)(t / cout);
      wf[i] = f2bf(pi);
      wb[((long long)tap * cin + ci) * cout + co] = f2bf(pi);
    }
  }
}

// a[1] = n / a[0]: the reciprocal of the mean from the sum
__global__ void k_recip_mean(float *a, float n) { a[1] = n / a[0]; }

unsigned grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65535LL * 8) b = 65535LL * 8;
  return (unsigned)(b < 1 ? 1 : b);
}

// ------------------------------------------------------------ the network

struct ConvSpec {
  int cin, cout;
};
// c1a c1b c2a c2b c3a c3b c4a c4b c5a c5b
constexpr ConvSpec kConv[10] = {{16, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64},
                                {64, 64}, {96, 32}, {32, 32}, {48, 16}, {16, 16}};
constexpr int kLevel[10] = {0, 0, 1, 1, 2, 2, 1, 1, 0, 0};

constexpr int kCoIB = 2;  // output planes per block of k_conv_co (32 channels)
constexpr int kInIB = 8;  // output planes per block of k_conv_in
template <int CIN, int COUT, int CA = 0, bool UP = false>
void conv_fwd_t(const bf16 *X, const bf16 *S, const bf16 *Wf, const float *b, bf16 *Y, Vol v,
                hipStream_t s) {
  if constexpr (COUT == 16)
    hipLaunchKernelGGL((k_conv_ip<CIN, CA, UP, false>),
                       dim3((v.W + kXT - 1) / kXT, (v.H + kYT - 1) / kYT, (v.D + kIB - 1) / kIB),
                       dim3(256), 0, s, X, S, Wf, b, Y, v, HeadArgs{});
  else if constexpr (COUT == 32)
    hipLaunchKernelGGL((k_conv_co<CIN, COUT, CA, UP, kCoIB>),
                       dim3((v.W + kXT - 1) / kXT, (v.H + kYT - 1) / kYT, (v.D + kCoIB - 1) / kCoIB),
                       dim3(256), 0, s, X, S, Wf, b, Y, v);
  else
    hipLaunchKernelGGL((k_conv_fwd<CIN, COUT, CA, UP>),
                       dim3((v.W + kXT - 1) / kXT, (v.H + kYT - 1) / kYT, v.D), dim3(256), 0, s,
                       X, S, Wf, b, Y, v);
}
// the fourth convolution (32 -> 32) with the second pooling in its epilogue
void conv_fwd_pool32(const bf16 *X, const bf16 *Wf, const float *b, bf16 *Y, bf16 *P, Vol v,
                     hipStream_t s) {
  hipLaunchKernelGGL((k_conv_co<32, 32, 0, false, 2, true>),
                     dim3((v.W + kXT - 1) / kXT, (v.H + kYT - 1) / kYT, (v.D + 1) / 2),
                     dim3(256), 0, s, X, nullptr, Wf, b, Y, v, P);
}
// the second convolution (16 -> 16) with the first pooling in its epilogue
void conv_fwd_pool(const bf16 *X, const bf16 *Wf, const float *b, bf16 *Y, bf16 *P, Vol v,
                   hipStream_t s) {
  hipLaunchKernelGGL((k_conv_ip<16, 0, false, false, true>),
                     dim3((v.W + kXT - 1) / kXT, (v.H + kYT - 1) / kYT, (v.D + kIB - 1) / kIB),
                     dim3(256), 0, s, X, nullptr, Wf, b, Y, v, HeadArgs{}, P);
}
// the last convolution (16 -> 16) with the head fused into its epilogue
void conv_fwd_head(const bf16 *X, const bf16 *Wf, const float *b, const HeadArgs &h, Vol v,
                   hipStream_t s) {
  hipLaunchKernelGGL((k_conv_ip<16, 0, false, true>),
                     dim3((v.W + kXT - 1) / kXT, (v.H + kYT - 1) / kYT, (v.D + kIB - 1) / kIB),
                     dim3(256), 0, s, X, nullptr, Wf, b, nullptr, v, h);
}
template <int CIN, int COUT>
void conv_bwd_t(const bf16 *G, const bf16 *Yo, const bf16 *X, const bf16 *Wb, bf16 *GX,
                float *P, float *Pb, Vol v, hipStream_t s) {
  if (GX)
    hipLaunchKernelGGL((k_conv_bwd_data<CIN, COUT>), dim3((v.W + 63) / 64, v.H, v.D),
                       dim3(256), 0, s, G, Yo, Wb, GX, v);
  const int nch = (int)((v.n() + kWChunk - 1) / kWChunk);
  hipLaunchKernelGGL((k_conv_bwd_w<CIN, COUT>), dim3(nch, 27), dim3(256), 0, s, G, Yo, X,
                     P, v);
  hipLaunchKernelGGL((k_bias_grad<COUT>), dim3(nch), dim3(256), 0, s, G, Yo, Pb, v.n());
}

#define CONV_SWITCH(I, CALL)                          \
  switch (I) {                                        \
    case 0: case 1: case 9: { constexpr int A = 16, B = 16; CALL; } break; \
    case 2: { constexpr int A = 16, B = 32; CALL; } break;                 \
    case 3: case 7: { constexpr int A = 32, B = 32; CALL; } break;         \
    case 4: { constexpr int A = 32, B = 64; CALL; } break;                 \
    case 5: { constexpr int A = 64, B = 64; CALL; } break;                 \
    case 6: { constexpr int A = 96, B = 32; CALL; } break;                 \
    case 8: { constexpr int A = 48, B = 16; CALL; } break;                 \
  }

}  // namespace

class UNet {
 public:
  // parameter blocks: 10 conv weights [27][co][ci], 10 biases, head (17)
  std::vector<long long> off_, len_;
  long long nparam_ = 0;
  DevBuf p_, m_, s_, g_;  // float32 masters, Adam moments, gradients
  DevBuf wf_[10], wb_[10];
  long long step_ = 0;
  // activation buffers for the current shape
  Vol v0_{0, 0, 0};
  bool train_ready_ = false;
  DevBuf in_, act_[10], pool1_, pool2_, cat2_, cat1_, out_, invm_, parts_;
  DevBuf gact_[10], gcat2_, gcat1_, gskip1_, gskip2_, gp1_, gp2_, gd1_, wpart_, bpart_;
  DevBuf lossp_, whp_, sumbuf_;

  ClarityResult_t init(unsigned seed) {
    long long o = 0;
    for (int i = 0; i < 10; i++) {
      off_.push_back(o);
      len_.push_back(27LL * kConv[i].cin * kConv[i].cout);
      o += len_.back();
    }
    for (int i = 0; i < 10; i++) {
      off_.push_back(o);
      len_.push_back(kConv[i].cout);
      o += kConv[i].cout;
    }
    off_.push_back(o);
    len_.push_back(17);
    o += 17;
    nparam_ = o;
    std::vector<float> h(nparam_, 0.f);
    std::mt19937 rng(seed);
    for (int i = 0; i < 10; i++) {
      // He initialisation over the real fan-in (layer 0 has 3 live channels)
      const int fan = 27 * (i == 0 ? 3 : kConv[i].cin);
      std::normal_distribution<float> nd(0.f, std::sqrt(2.f / fan));
      for (long long k = 0; k < len_[i]; k++) {
        const int ci = (int)(k % kConv[i].cin);
        h[off_[i] + k] = (i == 0 && ci >= 3) ? 0.f : nd(rng);
      }
    }
    // head starts near zero: the untrained network returns the observation
    std::normal_distribution<float> hd(0.f, 1e-3f);
    for (int c = 0; c < 16; c++) h[off_[20] + c] = hd(rng);
    CL_TRY(p_.alloc(sizeof(float) * nparam_));
    CL_TRY(m_.alloc(sizeof(float) * nparam_));
    CL_TRY(s_.alloc(sizeof(float) * nparam_));
    CL_TRY(g_.alloc(sizeof(float) * nparam_));
    for (int i = 0; i < 10; i++) {
      CL_TRY(wf_[i].alloc(sizeof(bf16) * len_[i]));
      CL_TRY(wb_[i].alloc(sizeof(bf16) * len_[i]));
    }
    return set_params(h.data());
  }

  ClarityResult_t set_params(const float *host) {
    hipStream_t s = rt().stream;
    CL_HIP(hipMemcpyAsync(p_.p, host, sizeof(float) * nparam_, hipMemcpyHostToDevice, s));
    CL_HIP(hipMemsetAsync(m_.p, 0, sizeof(float) * nparam_, s));
    CL_HIP(hipMemsetAsync(s_.p, 0, sizeof(float) * nparam_, s));
    step_ = 0;
    // bf16 copies: an Adam step with lr 0 rewrites them from the masters
    CL_HIP(hipMemsetAsync(g_.p, 0, sizeof(float) * nparam_, s));
    for (int i = 0; i < 10; i++)
      hipLaunchKernelGGL(k_adam, dim3(grid_for(len_[i])), dim3(256), 0, s, p_.f() + off_[i],
                         m_.f() + off_[i], s_.f() + off_[i], g_.f() + off_[i], len_[i], 0.f,
                         0.f, 0.f, 1.f, 1.f, 1.f, (bf16 *)wf_[i].p, (bf16 *)wb_[i].p,
                         kConv[i].cin, kConv[i].cout);
    CL_TRY(launch_check());
    CL_HIP(hipStreamSynchronize(s));
    return CLARITY_SUCCESS;
  }

  ClarityResult_t shape(Vol v, bool train) {
    if (v.D % 4 || v.H % 4 || v.W % 4 || v.D < 4 || v.H < 4 || v.W < 4)
      return CLARITY_INVALID_ARGUMENT;
    // k_conv_ip addresses a plane of up to 48 bfloat16 channels with 32-bit
    // byte offsets
    if ((long long)v.H * v.W * 48 * 2 >= (1LL << 32)) return CLARITY_INVALID_ARGUMENT;
    const bool same = v.D == v0_.D && v.H == v0_.H && v.W == v0_.W;
    if (same && (!train || train_ready_)) return CLARITY_SUCCESS;
    const long long n[3] = {v.n(), v.n() / 8, v.n() / 64};
    if (!same) {
      v0_ = Vol{0, 0, 0};  // stays invalid until every buffer is sized
      train_ready_ = false;
      CL_TRY(alloc_forward(n));
      v0_ = v;
    }
    if (train && !train_ready_) {
      CL_TRY(alloc_training(n));
      train_ready_ = true;
    }
    return CLARITY_SUCCESS;
  }

  ClarityResult_t alloc_forward(const long long n[3]) {
    CL_TRY(in_.alloc(sizeof(bf16) * n[0] * 16));
    for (int i = 0; i < 10; i++)
      CL_TRY(act_[i].alloc(sizeof(bf16) * n[kLevel[i]] * kConv[i].cout));
    CL_TRY(pool1_.alloc(sizeof(bf16) * n[1] * 16));
    CL_TRY(pool2_.alloc(sizeof(bf16) * n[2] * 32));
    CL_TRY(out_.alloc(sizeof(float) * n[0]));
    CL_TRY(invm_.alloc(sizeof(float) * 4));
    CL_TRY(parts_.alloc(sizeof(double) * kParts));
    return CLARITY_SUCCESS;
  }

  ClarityResult_t alloc_training(const long long n[3]) {
    {
      CL_TRY(cat2_.alloc(sizeof(bf16) * n[1] * 96));
      CL_TRY(cat1_.alloc(sizeof(bf16) * n[0] * 48));
      for (int i = 0; i < 10; i++)
        CL_TRY(gact_[i].alloc(sizeof(bf16) * n[kLevel[i]] * kConv[i].cout));
      CL_TRY(gcat2_.alloc(sizeof(bf16) * n[1] * 96));
      CL_TRY(gcat1_.alloc(sizeof(bf16) * n[0] * 48));
      CL_TRY(gskip1_.alloc(sizeof(bf16) * n[0] * 16));
      CL_TRY(gskip2_.alloc(sizeof(bf16) * n[1] * 32));
      CL_TRY(gp1_.alloc(sizeof(bf16) * n[1] * 16));
      CL_TRY(gp2_.alloc(sizeof(bf16) * n[2] * 32));
      CL_TRY(gd1_.alloc(sizeof(bf16) * n[0] * 16));
      const long long nch = (n[0] + kWChunk - 1) / kWChunk;
      CL_TRY(wpart_.alloc(sizeof(float) * nch * 27 * 64 * 64));
      CL_TRY(bpart_.alloc(sizeof(float) * nch * 64));
      CL_TRY(lossp_.alloc(sizeof(double) * 1024));
      CL_TRY(whp_.alloc(sizeof(float) * 1024 * 17));
      CL_TRY(sumbuf_.alloc(sizeof(float) * 32));
    }
    return CLARITY_SUCCESS;
  }

  // forward on a device observation; leaves activations for backward (and,
  // with keep_concat, the decoder's concatenated inputs).  With head_out
  // (inference) the network input is built inside the first convolution
  // and the head runs in the last one's epilogue, writing head_out; neither
  // the 16-channel input nor the last activation is stored.
  ClarityResult_t forward(const float *obs_d, float sxy, float sz, bool keep_concat,
                          hipStream_t s, float *head_out = nullptr, bool rescale = false) {
    const Vol v = v0_;
    const Vol v1{v.D / 2, v.H / 2, v.W / 2}, v2{v.D / 4, v.H / 4, v.W / 4};
    const Vol lv[3] = {v, v1, v2};
    // 1 / mean(obs)
    CL_TRY(k_sum_partials(obs_d, v.n(), parts_.d(), kParts, s));
    CL_TRY(k_finish_sum(parts_.d(), kParts, invm_.f(), nullptr, nullptr, s));
    hipLaunchKernelGGL(k_recip_mean, dim3(1), dim3(1), 0, s, invm_.f(), (float)v.n());
    if (!head_out)
      hipLaunchKernelGGL(k_input, dim3(grid_for(v.n())), dim3(256), 0, s, obs_d,
                         invm_.f() + 1, sxy, sz, (bf16 *)in_.p, v.n());
    const bf16 *src[10];
    src[0] = (bf16 *)in_.p;
    src[1] = (bf16 *)act_[0].p;
    src[2] = (bf16 *)pool1_.p;
    src[3] = (bf16 *)act_[2].p;
    src[4] = (bf16 *)pool2_.p;
    src[5] = (bf16 *)act_[4].p;
    src[6] = (bf16 *)cat2_.p;
    src[7] = (bf16 *)act_[6].p;
    src[8] = (bf16 *)cat1_.p;
    src[9] = (bf16 *)act_[8].p;
    for (int i = 0; i < 10; i++) {
      // the two poolings run in the epilogues of layers 1 and 3
      bf16 *Y = (bf16 *)act_[i].p;
      const float *b = p_.f() + off_[10 + i];
      const bf16 *Wf = (bf16 *)wf_[i].p;
      const Vol vv = lv[kLevel[i]];
      if (i == 1) {
        conv_fwd_pool(src[1], Wf, b, Y, (bf16 *)pool1_.p, v, s);
        continue;
      }
      if (i == 3) {
        static_assert(kCoIB == 2, "the fused pooling takes two planes per block");
        conv_fwd_pool32(src[3], Wf, b, Y, (bf16 *)pool2_.p, v1, s);
        continue;
      }
      if (i == 0 && head_out) {
        hipLaunchKernelGGL((k_conv_in<kInIB>),
                           dim3((v.W + kXT - 1) / kXT, (v.H + kYT - 1) / kYT,
                                (v.D + kInIB - 1) / kInIB),
                           dim3(256), 0, s, obs_d, invm_.f() + 1, sxy, sz, Wf, b, Y, v);
        continue;
      }
      if (i == 9 && head_out) {
        HeadArgs h;
        h.obs = obs_d;
        h.inv_mean = invm_.f() + 1;
        h.w = p_.f() + off_[20];
        h.out = head_out;
        h.rescale = rescale ? 1 : 0;
        conv_fwd_head((bf16 *)act_[8].p, Wf, b, h, v, s);
        continue;
      }
      // decoder inputs up(coarse) | skip: read in place for inference; the
      // training pass materialises them (the weight gradient reads them)
      if (i == 6 && !keep_concat) {
        conv_fwd_t<96, 32, 64, true>((bf16 *)act_[5].p, (bf16 *)act_[3].p, Wf, b, Y, v1, s);
        continue;
      }
      if (i == 8 && !keep_concat) {
        conv_fwd_t<48, 16, 32, true>((bf16 *)act_[7].p, (bf16 *)act_[1].p, Wf, b, Y, v, s);
        continue;
      }
      if (i == 6)
        hipLaunchKernelGGL(k_up_concat, dim3(grid_for(v1.n() * 96)), dim3(256), 0, s,
                           (bf16 *)act_[5].p, (bf16 *)act_[3].p, (bf16 *)cat2_.p, v1, 64, 96);
      if (i == 8)
        hipLaunchKernelGGL(k_up_concat, dim3(grid_for(v.n() * 48)), dim3(256), 0, s,
                           (bf16 *)act_[7].p, (bf16 *)act_[1].p, (bf16 *)cat1_.p, v, 32, 48);
      const bf16 *X = src[i];
      CONV_SWITCH(i, (conv_fwd_t<A, B>(X, nullptr, Wf, b, Y, vv, s)));
    }
    CL_TRY(launch_check());
    return CLARITY_SUCCESS;
  }

  ClarityResult_t head(const float *obs_d, float *out_d, bool rescale, hipStream_t s) {
    hipLaunchKernelGGL(k_head, dim3(grid_for(v0_.n())), dim3(256), 0, s,
                       (bf16 *)act_[9].p, p_.f() + off_[20], obs_d, invm_.f() + 1, out_d,
                       v0_.n(), rescale ? 1 : 0);
    return launch_check();
  }

  // loss and gradients for target tgt_d; Adam step when lr > 0
  ClarityResult_t backward(const float *tgt_d, float lr, float *loss_h, hipStream_t s) {
    const Vol v = v0_;
    const Vol v1{v.D / 2, v.H / 2, v.W / 2}, v2{v.D / 4, v.H / 4, v.W / 4};
    const Vol lv[3] = {v, v1, v2};
    CL_HIP(hipMemsetAsync(g_.p, 0, sizeof(float) * nparam_, s));
    const int nb = 1024;
    hipLaunchKernelGGL(k_loss_head_bwd, dim3(nb), dim3(256), 0, s, out_.f(), tgt_d,
                       invm_.f() + 1, (bf16 *)act_[9].p, p_.f() + off_[20],
                       (bf16 *)gd1_.p, lossp_.d(), whp_.f(), v.n());
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(64), 0, s, whp_.f(), nb, 17LL,
                       g_.f() + off_[20]);
    // input of each conv, gradient buffers of each conv's output
    const bf16 *xin[10] = {(bf16 *)in_.p,    (bf16 *)act_[0].p, (bf16 *)pool1_.p,
                           (bf16 *)act_[2].p, (bf16 *)pool2_.p,  (bf16 *)act_[4].p,
                           (bf16 *)cat2_.p,   (bf16 *)act_[6].p, (bf16 *)cat1_.p,
                           (bf16 *)act_[8].p};
    bf16 *gout[10];
    for (int i = 0; i < 9; i++) gout[i] = (bf16 *)gact_[i].p;
    gout[9] = (bf16 *)gd1_.p;
    bf16 *gin[10] = {nullptr,          gout[0], (bf16 *)gp1_.p, gout[2], (bf16 *)gp2_.p,
                     gout[4],          (bf16 *)gcat2_.p, gout[6], (bf16 *)gcat1_.p, gout[8]};
    for (int i = 9; i >= 0; i--) {
      const Vol vv = lv[kLevel[i]];
      const int nch = (int)((vv.n() + kWChunk - 1) / kWChunk);
      const bf16 *Wb = (bf16 *)wb_[i].p;
      CONV_SWITCH(i, (conv_bwd_t<A, B>(gout[i], (bf16 *)act_[i].p, xin[i], Wb, gin[i],
                                       wpart_.f(), bpart_.f(), vv, s)));
      hipLaunchKernelGGL(k_reduce_chunks, dim3(grid_for(len_[i])), dim3(256), 0, s,
                         wpart_.f(), nch, len_[i], g_.f() + off_[i]);
      hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(64), 0, s, bpart_.f(), nch,
                         (long long)kConv[i].cout, g_.f() + off_[10 + i]);
      // route gradients through concat / upsampling / pooling
      if (i == 8)  // cat1 = up(d2) | e1
        hipLaunchKernelGGL(k_up_concat_bwd, dim3(grid_for(v.n() * 32)), dim3(256), 0, s,
                           (bf16 *)gcat1_.p, gout[7], (bf16 *)gskip1_.p, v, 32, 48);
      if (i == 6)  // cat2 = up(b3) | e2
        hipLaunchKernelGGL(k_up_concat_bwd, dim3(grid_for(v1.n() * 64)), dim3(256), 0, s,
                           (bf16 *)gcat2_.p, gout[5], (bf16 *)gskip2_.p, v1, 64, 96);
      if (i == 4) {  // e2 feeds pool2 and the skip: sum both gradients
        CL_HIP(hipMemcpyAsync(gout[3], gskip2_.p, sizeof(bf16) * v1.n() * 32,
                              hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_pool_bwd_add, dim3(grid_for(v1.n() * 32)), dim3(256), 0, s,
                           (bf16 *)gp2_.p, gout[3], v1, 32);
      }
      if (i == 2) {
        CL_HIP(hipMemcpyAsync(gout[1], gskip1_.p, sizeof(bf16) * v.n() * 16,
                              hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_pool_bwd_add, dim3(grid_for(v.n() * 16)), dim3(256), 0, s,
                           (bf16 *)gp1_.p, gout[1], v, 16);
      }
    }
    CL_TRY(launch_check());
    std::vector<double> lp(nb);
    CL_HIP(hipMemcpyAsync(lp.data(), lossp_.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    if (lr > 0.f) {
      step_++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)step_), c2 = 1.f - std::pow(b2, (float)step_);
      for (int i = 0; i < 21; i++) {
        const int cin = i < 10 ? kConv[i].cin : 0, cout = i < 10 ? kConv[i].cout : 0;
        hipLaunchKernelGGL(k_adam, dim3(grid_for(len_[i])), dim3(256), 0, s, p_.f() + off_[i],
                           m_.f() + off_[i], s_.f() + off_[i], g_.f() + off_[i], len_[i], lr,
                           b1, b2, 1e-8f, c1, c2, i < 10 ? (bf16 *)wf_[i].p : nullptr,
                           i < 10 ? (bf16 *)wb_[i].p : nullptr, cin, cout);
      }
      CL_TRY(launch_check());
    }
    CL_HIP(hipStreamSynchronize(s));
    double l = 0.0;
    for (double x : lp) l += x;
    if (loss_h) *loss_h = (float)(l / (double)v.n());
    return CLARITY_SUCCESS;
  }
};

}  // namespace clarity

using namespace clarity;

struct ClarityEx_UNet_t {
  RegHold reg;  // first: released after the network's buffers
  UNet net;
};

namespace {

#define API_LOCK std::lock_guard<std::recursive_mutex> lk_(rt().mu)

ClarityResult_t unet_run(ClarityEx_UNet h, const float *obs, const float *tgt,
                         Clarity_Dim3 d, float sxy, float sz, float lr, float *out,
                         float *loss, float *grads, int on_device) {
  if (!h || !obs || !dims_ok(d)) return CLARITY_INVALID_ARGUMENT;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  UNet &n = h->net;
  const bool train = tgt != nullptr;
  const Vol v{d.z, d.y, d.x};
  CL_TRY(n.shape(v, train));
  hipStream_t s = rt().stream;
  const long long nv = v.n();
  DevBuf a, b, o;
  const float *obs_d = obs, *tgt_d = tgt;
  if (!on_device) {
    CL_TRY(a.alloc(sizeof(float) * nv));
    CL_HIP(hipMemcpyAsync(a.p, obs, sizeof(float) * nv, hipMemcpyHostToDevice, s));
    obs_d = a.f();
    if (train) {
      CL_TRY(b.alloc(sizeof(float) * nv));
      CL_HIP(hipMemcpyAsync(b.p, tgt, sizeof(float) * nv, hipMemcpyHostToDevice, s));
      tgt_d = b.f();
    }
  }
  if (train) {
    CL_TRY(n.forward(obs_d, sxy, sz, true, s));
    CL_TRY(n.head(obs_d, n.out_.f(), false, s));
    CL_TRY(n.backward(tgt_d, lr, loss, s));
    if (grads)
      CL_HIP(hipMemcpy(grads, n.g_.p, sizeof(float) * n.nparam_, hipMemcpyDeviceToHost));
    return CLARITY_SUCCESS;
  }
  if (on_device) {
    CL_TRY(n.forward(obs_d, sxy, sz, false, s, out, true));
  } else {
    CL_TRY(o.alloc(sizeof(float) * nv));
    CL_TRY(n.forward(obs_d, sxy, sz, false, s, o.f(), true));
    CL_HIP(hipMemcpyAsync(out, o.p, sizeof(float) * nv, hipMemcpyDeviceToHost, s));
  }
  CL_HIP(hipStreamSynchronize(s));
  return CLARITY_SUCCESS;
}

}  // namespace

extern "C" {

ClarityResult_t ClarityEx_UNetCreate(ClarityEx_UNet *net, unsigned seed) {
  API_LOCK;
  if (!net) return CLARITY_INVALID_ARGUMENT;
  *net = nullptr;
  if (!ready()) return CLARITY_INVALID_OPERATION;
  std::unique_ptr<ClarityEx_UNet_t> h(new ClarityEx_UNet_t);
  CL_TRY(h->reg.acquire());
  CL_TRY(h->net.init(seed));
  *net = h.release();
  return CLARITY_SUCCESS;
}

ClarityResult_t ClarityEx_UNetDestroy(ClarityEx_UNet net) {
  API_LOCK;
  if (!net) return CLARITY_INVALID_ARGUMENT;
  if (rt().stream) (void)hipStreamSynchronize(rt().stream);
  delete net;
  return CLARITY_SUCCESS;
}

int64_t ClarityEx_UNetNumParams(ClarityEx_UNet net) { return net ? net->net.nparam_ : 0; }

ClarityResult_t ClarityEx_UNetGetParams(ClarityEx_UNet net, float *params) {
  API_LOCK;
  if (!net || !params) return CLARITY_INVALID_ARGUMENT;
  CL_HIP(hipStreamSynchronize(rt().stream));
  CL_HIP(hipMemcpy(params, net->net.p_.p, sizeof(float) * net->net.nparam_,
                   hipMemcpyDeviceToHost));
  return CLARITY_SUCCESS;
}

ClarityResult_t ClarityEx_UNetSetParams(ClarityEx_UNet net, const float *params) {
  API_LOCK;
  if (!net || !params) return CLARITY_INVALID_ARGUMENT;
  return net->net.set_params(params);
}

ClarityResult_t ClarityEx_UNetPredict(ClarityEx_UNet net, const float *obs,
                                      Clarity_Dim3 dims, float sigmaXY, float sigmaZ,
                                      float *out, int onDevice) {
  API_LOCK;
  if (!out) return CLARITY_INVALID_ARGUMENT;
  return unet_run(net, obs, nullptr, dims, sigmaXY, sigmaZ, 0.f, out, nullptr, nullptr,
                  onDevice);
}

ClarityResult_t ClarityEx_UNetTrainStep(ClarityEx_UNet net, const float *obs,
                                        const float *target, Clarity_Dim3 dims,
                                        float sigmaXY, float sigmaZ, float learningRate,
                                        float *loss, float *grads, int onDevice) {
  API_LOCK;
  if (!target || learningRate < 0.f) return CLARITY_INVALID_ARGUMENT;
  return unet_run(net, obs, target, dims, sigmaXY, sigmaZ, learningRate, nullptr, loss,
                  grads, onDevice);
}

}  // extern "C"
