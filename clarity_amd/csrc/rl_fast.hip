// Fused, pruned Richardson-Lucy pipeline for work shapes whose axes are
// 2^k or 3 * 2^k (x 128..1024, y 128..2048, z 64..512).
//
// The RL estimate and the ratio obs/s1 are zero outside the image corner of
// the work domain (the estimate starts as the zero-padded observation and
// every update multiplies it, reference src/MaximumLikelihoodDeconvolve.cxx:
// 69-70,137), and s1, s2 are only ever read inside the corner.  So each 3D
// transform reads only the nonzero corner and each inverse writes only the
// corner: per RL iteration nine kernels stream
//
//   x_fwd   u (image rows)            -> S      R2C along x, rows y<iy, z<iz
//   y_fwd   S (y<iy)                  -> S      FFT along y, all ny outputs
//   z_mod   S (z<iz), K               -> S      FFT z of S and K, * /N, IFFT z, z<iz kept
//   y_inv   S                         -> S      IFFT along y, y<iy kept
//   x_mid   S, obs                    -> S      C2R x, r = guard(obs/(c*s1)), R2C x
//   y_fwd, z_mod, y_inv               (second convolution)
//   x_last  S, u                      -> u      C2R x, u' = (c*u)*s2, fp64 partials
//   finish  partials                  -> c' = E_obs / sum(u')
//
// with the estimate kept as u and its energy scale c in HBM (g = c*u), so no
// iteration waits on the host.  Between consecutive iterations x_last and
// the next x_fwd run as one kernel (x_last_fwd: u' is transformed while
// still in registers), so a run of n iterations launches x_fwd and x_last
// once and x_last_fwd n - 1 times.  The spectrum S has layout (z, kx, y), y
// fastest, z < iz only.  The kernel enters as K, its x and y transform
// before the z transform (the OTF's last step), laid out (p, kx, ky) over
// only the kz planes p the -k/2 shifted kernel occupies: z_mod finishes each
// OTF column along z in registers instead of reading a stored OTF of all nz
// planes (a quarter of its bytes at a same-size kernel, for one more FFT).
// Every transform is
// a register-resident Stockham FFT: a thread holds E elements of its pencil
// (element t + T*e in v[e]), butterflies of radix E run in registers and
// the stages exchange through LDS.  The x passes use the half-length complex
// FFT of (x[2m] + i x[2m+1]) with the usual pack/unpack.
#include "complex_ops.hpp"
#include "rl_plan.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace clarity {
namespace {

// Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format;
// no stride, swizzle or index).
constexpr int kBufWord3 = 0x00020000;

// ---------------------------------------------------------------- butterflies

// cos / sin of 2*pi*j/16
__device__ constexpr float kC16[16] = {
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f,
    0.0f, -0.38268343236508978f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508978f,
    0.0f, 0.38268343236508978f, 0.70710678118654752f, 0.92387953251128674f};
__device__ constexpr float kS16[16] = {
    0.0f, 0.38268343236508978f, 0.70710678118654752f, 0.92387953251128674f,
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f,
    0.0f, -0.38268343236508978f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508978f};

// a * exp(SIGN * 2 pi i j / 16) for a compile-time j
template <int SIGN, int J>
__device__ __forceinline__ float2 rot16(float2 a) {
  constexpr int j = J & 15;
  if constexpr (j == 0) {
    return a;
  } else if constexpr (j == 4) {
    return cmul_si(a, SIGN);
  } else if constexpr (j == 8) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (j == 12) {
    return cmul_si(a, -SIGN);
  } else {
    const float c = kC16[j];
    const float s = SIGN * kS16[j];
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// 16-point DFT: n = 4*n1 + n2, k = k1 + 4*k2; four DFT4 over n1, twiddle
// w16^(n2*k1), four DFT4 over n2.
template <int SIGN>
__device__ __forceinline__ void dft16(float2 (&u)[16]) {
  float2 a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) {
    float2 t[4] = {u[n2], u[4 + n2], u[8 + n2], u[12 + n2]};
    dft4(t, SIGN);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) a[n2][k1] = t[k1];
  }
  a[1][1] = rot16<SIGN, 1>(a[1][1]);
  a[1][2] = rot16<SIGN, 2>(a[1][2]);
  a[1][3] = rot16<SIGN, 3>(a[1][3]);
  a[2][1] = rot16<SIGN, 2>(a[2][1]);
  a[2][2] = rot16<SIGN, 4>(a[2][2]);
  a[2][3] = rot16<SIGN, 6>(a[2][3]);
  a[3][1] = rot16<SIGN, 3>(a[3][1]);
  a[3][2] = rot16<SIGN, 6>(a[3][2]);
  a[3][3] = rot16<SIGN, 9>(a[3][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    float2 t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4(t, SIGN);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) u[k1 + 4 * k2] = t[k2];
  }
}

// 3-point DFT: X1,2 = a - (b+c)/2 +- SIGN*i*(sqrt(3)/2)*(b-c)
template <int SIGN>
__device__ __forceinline__ void dft3(float2 &a, float2 &b, float2 &c) {
  constexpr float h = 0.86602540378443864676f;
  const float2 t1 = cadd(b, c);
  const float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
  const float2 d = csub(b, c);
  const float2 r = make_float2(-SIGN * h * d.y, SIGN * h * d.x);  // SIGN*i*h*d
  a = cadd(a, t1);
  b = cadd(t2, r);
  c = csub(t2, r);
}

// 6-point DFT: n = 3*n1 + n2, k = k1 + 2*k2; three DFT2 over n1, twiddle
// w6^(n2*k1), two DFT3 over n2.
template <int SIGN>
__device__ __forceinline__ void dft6(float2 (&u)[6]) {
  constexpr float h = 0.86602540378443864676f;
  float2 a0 = u[0], b0 = u[3], a1 = u[1], b1 = u[4], a2 = u[2], b2 = u[5];
  dft2(a0, b0);
  dft2(a1, b1);
  dft2(a2, b2);
  // w6^1 = 1/2 + SIGN*i*h, w6^2 = -1/2 + SIGN*i*h
  b1 = make_float2(0.5f * b1.x - SIGN * h * b1.y, SIGN * h * b1.x + 0.5f * b1.y);
  b2 = make_float2(-0.5f * b2.x - SIGN * h * b2.y, SIGN * h * b2.x - 0.5f * b2.y);
  dft3<SIGN>(a0, a1, a2);
  dft3<SIGN>(b0, b1, b2);
  u[0] = a0; u[2] = a1; u[4] = a2;
  u[1] = b0; u[3] = b1; u[5] = b2;
}

// DFT4 with zero inputs Z (bit n: u[n] = 0) and outputs k < KN only.
template <int Z, int KN>
__device__ __forceinline__ void dft4_p(float2 (&u)[4], int sign) {
  constexpr bool z1 = Z & 2, z2 = Z & 4, z3 = Z & 8;
  const float2 t0 = z2 ? u[0] : cadd(u[0], u[2]);
  const float2 t1 = z2 ? u[0] : csub(u[0], u[2]);
  const float2 t2 = z3 ? u[1] : z1 ? u[3] : cadd(u[1], u[3]);
  const float2 d = z3 ? u[1] : z1 ? make_float2(-u[3].x, -u[3].y) : csub(u[1], u[3]);
  const float2 t3 = cmul_si(d, sign);
  u[0] = cadd(t0, t2);
  u[1] = cadd(t1, t3);
  if constexpr (KN > 2) u[2] = csub(t0, t2);
  if constexpr (KN > 3) u[3] = csub(t1, t3);
}

// a * exp(SIGN * 2 pi i j / 12) for a compile-time j
template <int SIGN, int J>
__device__ __forceinline__ float2 rot12(float2 a) {
  constexpr int j = J % 12;
  constexpr float h = 0.86602540378443864676f;  // sin(pi/3)
  if constexpr (j == 0) {
    return a;
  } else if constexpr (j == 3) {
    return cmul_si(a, SIGN);
  } else if constexpr (j == 6) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (j == 9) {
    return cmul_si(a, -SIGN);
  } else {
    // cos and sin of 2 pi j / 12 for j = 1, 2, 4, 5, 7, 8, 10, 11
    constexpr float c = (j == 1 || j == 11) ? h : (j == 2 || j == 10) ? 0.5f
                        : (j == 4 || j == 8) ? -0.5f : -h;
    constexpr float sn = (j == 1 || j == 5) ? 0.5f : (j == 2 || j == 4) ? h
                         : (j == 7 || j == 11) ? -0.5f : -h;
    constexpr float sg = SIGN * sn;
    return make_float2(a.x * c - a.y * sg, a.x * sg + a.y * c);
  }
}

// 12-point DFT: n = 3*n1 + n2, k = k1 + 4*k2; three DFT4 over n1, twiddle
// w12^(n2*k1), four DFT3 over n2.  Z (bit n1: inputs with that n1 are
// zero) prunes the DFT4s as dft4_p.
template <int SIGN, int Z = 0>
__device__ __forceinline__ void dft12(float2 (&u)[12]) {
  float2 a[3][4];
#pragma unroll
  for (int n2 = 0; n2 < 3; n2++) {
    float2 t[4] = {u[n2], u[3 + n2], u[6 + n2], u[9 + n2]};
    dft4_p<Z, 4>(t, SIGN);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) a[n2][k1] = t[k1];
  }
  a[1][1] = rot12<SIGN, 1>(a[1][1]);
  a[1][2] = rot12<SIGN, 2>(a[1][2]);
  a[1][3] = rot12<SIGN, 3>(a[1][3]);
  a[2][1] = rot12<SIGN, 2>(a[2][1]);
  a[2][2] = rot12<SIGN, 4>(a[2][2]);
  a[2][3] = rot12<SIGN, 6>(a[2][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    float2 b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1];
    dft3<SIGN>(b0, b1, b2);
    u[k1] = b0;
    u[k1 + 4] = b1;
    u[k1 + 8] = b2;
  }
}

template <int R, int SIGN>
__device__ __forceinline__ void dft(float2 (&u)[R]) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    dft2(u[0], u[1]);
  } else if constexpr (R == 4) {
    dft4(u, SIGN);
  } else if constexpr (R == 8) {
    dft8(u, SIGN);
  } else if constexpr (R == 16) {
    dft16<SIGN>(u);
  } else if constexpr (R == 3) {
    dft3<SIGN>(u[0], u[1], u[2]);
  } else if constexpr (R == 6) {
    dft6<SIGN>(u);
  } else if constexpr (R == 12) {
    dft12<SIGN>(u);
  } else {
    static_assert(R == 16, "unsupported radix");
  }
}

// Pruned butterflies for transforms whose input is known to be zero in part
// or whose output is wanted only in part (the same-size z pass: the
// spectrum column is zero in its upper half, the kernel column in its
// middle half, and the inverse is stored for its lower half only).  ZP: 0
// none, 1 inputs r >= R/2 are zero, 2 inputs R/4 <= r < 3R/4 are zero.
// KP: 0 every output, 1 outputs r < R/2 only (the others are left unset).
// Terms known to be zero are left out, never added (x + 0 is not foldable
// in IEEE arithmetic, -0 + 0 being +0).
constexpr int kZeroUpper = 1, kZeroMiddle = 2, kKeepLower = 1;


template <int SIGN, int ZP, int KP>
__device__ __forceinline__ void dft16_p(float2 (&u)[16]) {
  // n = 4*n1 + n2: upper half zero is n1 >= 2, middle half zero n1 in {1, 2}
  constexpr int Z = ZP == kZeroUpper ? 0xC : ZP == kZeroMiddle ? 0x6 : 0;
  // k = k1 + 4*k2 < 8 is k2 < 2
  constexpr int KN = KP == kKeepLower ? 2 : 4;
  float2 a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) {
    float2 t[4] = {u[n2], u[4 + n2], u[8 + n2], u[12 + n2]};
    dft4_p<Z, 4>(t, SIGN);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) a[n2][k1] = t[k1];
  }
  a[1][1] = rot16<SIGN, 1>(a[1][1]);
  a[1][2] = rot16<SIGN, 2>(a[1][2]);
  a[1][3] = rot16<SIGN, 3>(a[1][3]);
  a[2][1] = rot16<SIGN, 2>(a[2][1]);
  a[2][2] = rot16<SIGN, 4>(a[2][2]);
  a[2][3] = rot16<SIGN, 6>(a[2][3]);
  a[3][1] = rot16<SIGN, 3>(a[3][1]);
  a[3][2] = rot16<SIGN, 6>(a[3][2]);
  a[3][3] = rot16<SIGN, 9>(a[3][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    float2 t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4_p<0, KN>(t, SIGN);
#pragma unroll
    for (int k2 = 0; k2 < KN; k2++) u[k1 + 4 * k2] = t[k2];
  }
}

template <int SIGN, int ZP, int KP>
__device__ __forceinline__ void dft8_p(float2 (&u)[8]) {
  // even inputs u[0,2,4,6] and odd u[1,3,5,7], each as a DFT4 over m (u[2m
  // (+1)]): upper half zero is m >= 2, middle half zero m in {1, 2}
  constexpr int Z = ZP == kZeroUpper ? 0xC : ZP == kZeroMiddle ? 0x6 : 0;
  constexpr float c = 0.70710678118654752440f;
  float2 e[4] = {u[0], u[2], u[4], u[6]};
  float2 o[4] = {u[1], u[3], u[5], u[7]};
  dft4_p<Z, 4>(e, SIGN);
  dft4_p<Z, 4>(o, SIGN);
  constexpr float s = (float)SIGN;
  const float2 o1 = make_float2(c * (o[1].x - s * o[1].y), c * (o[1].y + s * o[1].x));
  const float2 o2 = cmul_si(o[2], SIGN);
  const float2 o3 = make_float2(c * (-o[3].x - s * o[3].y), c * (-o[3].y + s * o[3].x));
  u[0] = cadd(e[0], o[0]);
  u[1] = cadd(e[1], o1);
  u[2] = cadd(e[2], o2);
  u[3] = cadd(e[3], o3);
  if constexpr (KP != kKeepLower) {
    u[4] = csub(e[0], o[0]);
    u[5] = csub(e[1], o1);
    u[6] = csub(e[2], o2);
    u[7] = csub(e[3], o3);
  }
}

// dft<R, SIGN> with the patterns applied where the radix has a pruned form
// (16, 8, and 2 for KP); other radices compute every term.
template <int R, int SIGN, int ZP, int KP>
__device__ __forceinline__ void dft_p(float2 (&u)[R]) {
  if constexpr (ZP == 0 && KP == 0) {
    dft<R, SIGN>(u);
  } else if constexpr (R == 16) {
    dft16_p<SIGN, ZP, KP>(u);
  } else if constexpr (R == 8) {
    dft8_p<SIGN, ZP, KP>(u);
  } else if constexpr (R == 12) {
    // n = 3*n1 + n2: upper half zero is n1 >= 2, middle half zero n1 in {1, 2}
    dft12<SIGN, ZP == kZeroUpper ? 0xC : ZP == kZeroMiddle ? 0x6 : 0>(u);
  } else if constexpr (R == 2 && ZP == 0) {
    u[0] = cadd(u[0], u[1]);
  } else {
    dft<R, SIGN>(u);
  }
}

// Power-of-two butterflies, radix-2 decimation in time, for the stages
// whose inputs are multiplied first (a twiddled Stockham stage, or the
// spectrum times the kernel's spectrum).  A twiddled radix-2 butterfly
// y0 = a + w b, y1 = a - w b takes four fused multiply-adds for y0 and two
// for y1 = 2 a - y0 (eight operations multiply-then-add), and an input
// multiplier folds into the first layer the same way: (m_a a) + m_b b is
// four more FMAs after the product m_a a.  A 16-point transform with its 15
// input twiddles drops from 220 to 192 VALU operations, an untwiddled one
// from 160 to 148.  y1 carries y0's rounding once more, an error of the
// same order as the butterfly's own.

// a + w b (a = y0 of the butterfly), w = (c, s)
__device__ __forceinline__ float2 fma_wb(float2 a, float2 b, float c, float s) {
  return make_float2(fmaf(-s, b.y, fmaf(c, b.x, a.x)), fmaf(s, b.x, fmaf(c, b.y, a.y)));
}
// 2 a - y0
__device__ __forceinline__ float2 twice_minus(float2 a, float2 y0) {
  return make_float2(fmaf(2.f, a.x, -y0.x), fmaf(2.f, a.y, -y0.y));
}

// Multipliers of a butterfly's inputs: kMulNone; kMulTw: m[r] for r >= 1
// (m[0] = 1, a twiddled stage); kMulAll: m[r] for every r (the product
// with the kernel's spectrum); kMulConj: conj(m[r]) for every r (a
// correlation).
constexpr int kMulNone = 0, kMulTw = 1, kMulAll = 2, kMulConj = 3;

template <int MUL>
__device__ __forceinline__ bool mul_unit(int r) {
  return MUL == kMulNone || (MUL == kMulTw && r == 0);
}
template <int MUL>
__device__ __forceinline__ float2 mul_w(const float2 &m) {
  return MUL == kMulConj ? make_float2(m.x, -m.y) : m;
}

// R-point DFT (R = 2, 4, 8, 16) of u[r] * m[r] (MUL), the zero patterns of
// dft_p on the inputs (ZP) and its kept outputs (KP); natural order out.
template <int R, int SIGN, int ZP, int KP, int MUL>
__device__ __forceinline__ void dft_dit(float2 (&u)[R], const float2 (&m)[R]) {
  constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
  static_assert((1 << LR) == R, "radix-2^k only");
  auto zero = [](int r) {
    return (ZP == kZeroUpper && r >= R / 2) ||
           (ZP == kZeroMiddle && r >= R / 4 && r < 3 * R / 4);
  };
  float2 a[R];
  // layer 1: a[2j], a[2j+1] from inputs n = bitrev(j), n + R/2
#pragma unroll
  for (int j = 0; j < R / 2; j++) {
    int n = 0;
#pragma unroll
    for (int b = 0; b < LR - 1; b++) n |= ((j >> b) & 1) << (LR - 2 - b);
    const int n1 = n + R / 2;
    if (zero(n1)) {
      const float2 x = mul_unit<MUL>(n) ? u[n] : cmul(u[n], mul_w<MUL>(m[n]));
      a[2 * j] = x;
      a[2 * j + 1] = x;
    } else if (zero(n)) {
      const float2 y = mul_unit<MUL>(n1) ? u[n1] : cmul(u[n1], mul_w<MUL>(m[n1]));
      a[2 * j] = y;
      a[2 * j + 1] = make_float2(-y.x, -y.y);
    } else if (MUL == kMulNone) {
      a[2 * j] = cadd(u[n], u[n1]);
      a[2 * j + 1] = csub(u[n], u[n1]);
    } else {
      const float2 x = mul_unit<MUL>(n) ? u[n] : cmul(u[n], mul_w<MUL>(m[n]));
      const float2 w = mul_w<MUL>(m[n1]);
      const float2 y0 = fma_wb(x, u[n1], w.x, w.y);
      a[2 * j] = y0;
      a[2 * j + 1] = twice_minus(x, y0);
    }
  }
  // layers of span h = 2, 4, .., R/2: pairs (g + k, g + k + h), twiddle
  // w_{2h}^k = w_16^(k * 8 / h)
#pragma unroll
  for (int h = 2; h < R; h *= 2) {
    const bool last = 2 * h == R;
#pragma unroll
    for (int g = 0; g < R; g += 2 * h) {
#pragma unroll
      for (int k = 0; k < h; k++) {
        const int J = k * 8 / h;  // in sixteenths of a turn
        const int i0 = g + k, i1 = i0 + h;
        const bool keep1 = !(last && KP == kKeepLower);
        if (J == 0) {
          const float2 y0 = cadd(a[i0], a[i1]);
          if (keep1) a[i1] = csub(a[i0], a[i1]);
          a[i0] = y0;
        } else if (J == 4) {
          const float2 b = cmul_si(a[i1], SIGN);
          if (keep1) a[i1] = csub(a[i0], b);
          a[i0] = cadd(a[i0], b);
        } else {
          const float2 y0 = fma_wb(a[i0], a[i1], kC16[J], SIGN * kS16[J]);
          if (keep1) a[i1] = twice_minus(a[i0], y0);
          a[i0] = y0;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++)
    if (!(KP == kKeepLower && r >= R / 2)) u[r] = a[r];
}

template <int SIGN>
__device__ __forceinline__ float2 tw_load(const float2 *__restrict__ tw, int idx) {
  float2 w = tw[idx];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// ------------------------------------------------------------- LDS layouts

// LDS exchanges between the lanes of one wavefront: the lanes run in
// lockstep and a wavefront's LDS operations complete in order, so a fence
// that keeps the compiler from reordering them is all that is needed.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS slot of element pos of an FFT line stored with one pad slot after
// every 2^PS elements.  A Stockham stage of radix R <= 2^PS writes its
// outputs at stride R; unpadded, the lanes of such a write pile onto a few
// banks, padded they spread over all of them.
template <int PS>
__device__ __forceinline__ int padded(int pos) {
  return pos + (pos >> PS);
}
template <int N, int PS>
constexpr int kPadded = N + (N >> PS);
// pad interval for E elements per thread (the first stage's radix): a slot
// every 16 elements at 16 per thread, every 4 at 12 per thread (every 8:
// x_mid 107.5 and x_last_fwd 148.6 us at 512x512x128 against 105.2 and
// 145.4; every 16 or 32 slower still; at 16 per thread every 8 or 4 took
// the 512-point x_mid from 133 to 204-215 us)
template <int E>
constexpr int kPadShift = (E == 16) ? 4 : 2;

// Row q of an x-pass tile: padded, and rows kRowStride apart, an odd number
// of slots, so that a column access (same pos, consecutive rows) also hits
// distinct banks.  Positions stay affine in the element index, so a lane
// addresses all its elements from one register plus immediate offsets.  A
// row is transformed by T <= 64 consecutive lanes, i.e. inside one
// wavefront.
template <int M, int PS>
constexpr int kRowStride = kPadded<M, PS> + 1;

template <int PS>
struct RowLds {
  float2 *base;
  __device__ __forceinline__ float2 &operator()(int pos) const { return base[padded<PS>(pos)]; }
  __device__ __forceinline__ void sync() const { wave_sync(); }
};
// Pencil q of an interleaved tile: element pos at pos*TP + q.  A pencil's
// threads are spread over the workgroup.
template <int TP>
struct InterLds {
  float2 *base;
  __device__ __forceinline__ float2 &operator()(int pos) const { return base[pos * TP]; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};
// A contiguous (padded) pencil transformed by T consecutive lanes; WAVE when
// T <= 64.
template <bool WAVE, int PS>
struct PencilLds {
  float2 *base;
  __device__ __forceinline__ float2 &operator()(int pos) const { return base[padded<PS>(pos)]; }
  __device__ __forceinline__ void sync() const {
    if constexpr (WAVE)
      wave_sync();
    else
      __syncthreads();
  }
};

// ------------------------------------------------------------ FFT engine

// Compile-time radix sequences.  The default for a power of two is E, E,
// ..., then the remainder; mixed sizes list theirs explicitly.
template <int... Rs>
struct Radices {};

template <int N, int E, int P, bool Done, int... Acc>
struct RadixList;
template <int N, int E, int P, int... Acc>
struct RadixList<N, E, P, true, Acc...> {
  using type = Radices<Acc...>;
};
template <int N, int E, int P, int... Acc>
struct RadixList<N, E, P, false, Acc...> {
  static constexpr int R = (N / P >= E) ? E : N / P;
  using type = typename RadixList<N, E, P * R, (P * R == N), Acc..., R>::type;
};
template <int N, int E>
using DefaultRadices = typename RadixList<N, E, 1, false>::type;

// Stockham stages from span P on, one per radix of the list.  Thread t of a
// pencil holds x[t + T*e] in v[e] on entry; on exit v[e] holds X[t + T*e]
// (natural order).  Stages exchange through the accessor's LDS and
// synchronise with its sync(): a wavefront-local barrier when every pencil
// lies inside one wavefront, else a workgroup barrier (then all threads of
// the workgroup must call this together).
//
// fft_pruned: the same with the pruning patterns of dft_p, ZP on the first
// stage's inputs and KP on the last stage's outputs, stated for v: ZP 1
// is v[e] = 0 for e >= E/2, ZP 2 for E/4 <= e < 3E/4; KP 1 asks for v[e],
// e < E/2, only (the rest of v is left unset).  With u[r] = v[b + r*NB]
// (NB = E/R, b < NB) these are the patterns r >= R/2, R/4 <= r < 3R/4 and
// r < R/2 of every butterfly of the stage.
template <int N, int E, int P, int SIGN, int ZP, int KP, int FM, class L, int R, int... Rest>
__device__ __forceinline__ void fft_pruned(float2 (&v)[E], int t, const L &lds,
                                           const float2 *__restrict__ tw,
                                           Radices<R, Rest...>, const float2 (&fm)[E]) {
  constexpr int T = N / E;
  constexpr bool LAST = sizeof...(Rest) == 0;
  constexpr int NB = E / R;
  // patterns that do not fall on whole butterfly inputs are not applied
  constexpr int ZS = (P == 1 && ((ZP == kZeroUpper && R % 2 == 0) ||
                                 (ZP == kZeroMiddle && R % 4 == 0)))
                         ? ZP
                         : 0;
  constexpr int KS = (LAST && KP == kKeepLower && R % 2 == 0) ? KP : 0;
  static_assert(E % R == 0, "radix must divide the per-thread count");
  static_assert(!LAST || P * R == N, "radices must multiply to N");
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int i = t + T * b;
    const int k = i % P;
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; r++) u[r] = v[b + r * NB];
    constexpr bool DIT = R == 2 || R == 4 || R == 8 || R == 16;
    if constexpr (DIT && (P > 1 || FM != kMulNone)) {
      // multipliers folded into the first butterfly layer
      float2 m[R];
      if constexpr (P > 1) {
        if constexpr (R == 4) {
          // one table read and powers by complex products (a few ulp more
          // error per twiddle, two LDS reads fewer per butterfly: the
          // 3 * 2^k row passes are bound by LDS instructions)
          m[1] = tw_load<SIGN>(tw, k * (N / (P * R)));
          m[2] = cmul(m[1], m[1]);
          m[3] = cmul(m[2], m[1]);
        } else {
#pragma unroll
          for (int r = 1; r < R; r++) m[r] = tw_load<SIGN>(tw, r * k * (N / (P * R)));
        }
        dft_dit<R, SIGN, ZS, KS, kMulTw>(u, m);
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) m[r] = fm[b + r * NB];
        dft_dit<R, SIGN, ZS, KS, FM>(u, m);
      }
    } else {
      if constexpr (FM != kMulNone) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const float2 w = fm[b + r * NB];
          u[r] = cmul(u[r], FM == kMulConj ? make_float2(w.x, -w.y) : w);
        }
      }
      if constexpr (P > 1) {
        if constexpr (R == 3) {
          const float2 w1 = tw_load<SIGN>(tw, k * (N / (P * R)));
          u[1] = cmul(u[1], w1);
          u[2] = cmul(u[2], cmul(w1, w1));
        } else {
#pragma unroll
          for (int r = 1; r < R; r++)
            u[r] = cmul(u[r], tw_load<SIGN>(tw, r * k * (N / (P * R))));
        }
      }
      dft_p<R, SIGN, ZS, KS>(u);
    }
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < (KS ? R / 2 : R); r++) v[b + r * NB] = u[r];
    } else {
      const int base = (i - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) lds(base + r * P) = u[r];
    }
  }
  if constexpr (!LAST) {
    lds.sync();
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = lds(t + T * e);
    lds.sync();
    fft_pruned<N, E, P * R, SIGN, 0, KP, kMulNone>(v, t, lds, tw, Radices<Rest...>{}, fm);
  }
}

// without input multipliers
template <int N, int E, int P, int SIGN, int ZP, int KP, class L, class Rad>
__device__ __forceinline__ void fft_pruned(float2 (&v)[E], int t, const L &lds,
                                           const float2 *__restrict__ tw, Rad r) {
  fft_pruned<N, E, P, SIGN, ZP, KP, kMulNone>(v, t, lds, tw, r, v);
}

template <int N, int E, int P, int SIGN, class L, class Rad>
__device__ __forceinline__ void fft_stages(float2 (&v)[E], int t, const L &lds,
                                           const float2 *__restrict__ tw, Rad r) {
  fft_pruned<N, E, P, SIGN, 0, 0>(v, t, lds, tw, r);
}

// Radix plan for an N-point transform with E elements per thread: powers of
// two use radix-E stages then the remainder; 3 * 2^k sizes use E = 12 with a
// radix-12 stage first (one LDS exchange fewer than radix 3 then 4s), but
// 384, where radix 12 would leave 32 = 4 * 4 * 2 and no stage is saved.
template <int N, int E>
struct Plan {
  using Rad = DefaultRadices<N, E>;
};
template <> struct Plan<96, 12> { using Rad = Radices<12, 4, 2>; };
template <> struct Plan<192, 12> { using Rad = Radices<12, 4, 4>; };
template <> struct Plan<384, 12> { using Rad = Radices<6, 4, 4, 4>; };
template <> struct Plan<768, 12> { using Rad = Radices<12, 4, 4, 4>; };

// Radix plans of the x passes, whose rows sit in LDS with one pad slot per
// 4 elements (12 per thread; 16 at 16 per thread): a stage writes its
// outputs at positions base + r * P (span P so far, r < R), which stay
// affine in r after padding -- one address register plus immediate
// offsets -- only when P is a multiple of the pad period or P * R stays
// inside one pad period (P = 1 with R <= 4).  A
// radix 6 or 12 stage first breaks that for the first two stages (3-4
// integer instructions per element written); taking it last instead costs
// its R - 1 twiddle products per butterfly and leaves every write affine:
// 384-point rows 1543 -> 1384 VALU instructions per wave in x_mid
// (113-116 -> 109 us at 512x512x128), 192-point rows 1338 -> 1240.  (At 96
// points, 8 lanes a row, it changes nothing.)  The y and z passes keep Plan (their
// LDS is unpadded, so radix 6 or 12 first costs no addressing and saves
// the twiddles).
template <int M, int E>
struct XPlan {
  using Rad = typename Plan<M, E>::Rad;
};
template <> struct XPlan<192, 12> { using Rad = Radices<4, 4, 12>; };
template <> struct XPlan<384, 12> { using Rad = Radices<4, 4, 4, 6>; };

// Entries of an N-point twiddle table that fft_pruned reads for a radix
// list: a stage of radix R at span P reads entry k * N / (P R) (R = 3, 4:
// one read, powers by products) or r k N / (P R), r < R (other radices),
// for k < P.  A kernel stages only these in LDS (a 768-point y pencil with
// radices 12, 4, 4, 4 reads the first 192 of 768).
template <int N, int P, class Rad>
struct TwNeed;
template <int N, int P>
struct TwNeed<N, P, Radices<>> {
  static constexpr int value = 1;
};
template <int N, int P, int R, int... Rest>
struct TwNeed<N, P, Radices<R, Rest...>> {
  static constexpr int here =
      P == 1 ? 1 : ((R == 3 || R == 4) ? (P - 1) : (R - 1) * (P - 1)) * (N / (P * R)) + 1;
  static constexpr int rest = TwNeed<N, P * R, Radices<Rest...>>::value;
  static constexpr int value = here > rest ? here : rest;
};
template <int N, int E>
constexpr int kTwNeed = TwNeed<N, 1, typename Plan<N, E>::Rad>::value;
// the largest index read plus one, from the stage loop run out by hand
static_assert(kTwNeed<768, 12> == 192 && kTwNeed<384, 12> == 96 && kTwNeed<192, 12> == 48 &&
                  kTwNeed<96, 12> == 48 && kTwNeed<1024, 16> == 901 &&
                  kTwNeed<512, 16> == 451 && kTwNeed<256, 16> == 226 &&
                  kTwNeed<128, 16> == 106 && kTwNeed<2048, 16> == 1801,
              "twiddle entries read");

struct Geom {
  int ix, iy, iz;  // image
  int nx, ny, nz;  // work
  int hx;          // nx/2 + 1
  long long sz;    // S / K stride of z: hx*ny
  // kernel planes: the -kz/2 shifted kernel occupies work planes z < ka
  // and z >= nz - (kz - ka), stored in K as planes p = z and
  // p = z - (nz - kz) (ka = kz - kz/2)
  int kz, ka;
  // x passes take their tiles from the last plane down (1) instead of up
  int xrev;
};

// An engine's device buffers and twiddle tables.
struct Bufs {
  float *u, *obs, *scal;  // scal: E, S, c
  float2 *S, *K;          // spectrum; kernel after its x and y transforms
  double *parts;
  const float2 *twM, *twN, *twY, *twZ;
};

// One volume of a batched launch: a kernel instantiation run over several
// volumes at once reads its volume's geometry and buffers here (the entry
// volume_of finds for blockIdx.x) and works on block blockIdx.x - blk0 of
// it.  The
// buffers' addresses in b are byte offsets from the table's own device
// address (relative_to, bufs_at): a pointer read from memory is generic to
// the compiler, which then addresses it with flat instructions (counted
// together with LDS operations, so every wait for an LDS access also waits
// for the memory accesses in flight); an address formed from the table's
// (a kernel argument, known to be global memory) gets global instructions.
struct VolTab {
  Geom g;
  Bufs b;
  int blk0;    // the volume's first block in the launch
  int nparts;  // finish: partial sums of the volume's last x pass
};

// The volume whose blocks include block b: the last i < n with blk0s[i] <= b
// (blk0s ascending, blk0s[0] = 0).  A binary search over the launch's short
// list of first blocks, which every workgroup reads (so it stays in the
// scalar cache), instead of a per-block map, whose entry for each new
// workgroup is a cache miss on the way to its first loads.
__device__ __forceinline__ int volume_of(const int *__restrict__ blk0s, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0s[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <class T>
T *offset_from(T *p, const void *base) {
  return reinterpret_cast<T *>(reinterpret_cast<uintptr_t>(p) -
                               reinterpret_cast<uintptr_t>(base));
}
// The table entry t with its buffer addresses made relative to base, the
// device address of the table it is copied to.
VolTab relative_to(VolTab t, const void *base) {
  Bufs &b = t.b;
  b.u = offset_from(b.u, base);
  b.obs = offset_from(b.obs, base);
  b.scal = offset_from(b.scal, base);
  b.S = offset_from(b.S, base);
  b.K = offset_from(b.K, base);
  b.parts = offset_from(b.parts, base);
  b.twM = offset_from(b.twM, base);
  b.twN = offset_from(b.twN, base);
  b.twY = offset_from(b.twY, base);
  b.twZ = offset_from(b.twZ, base);
  return t;
}
template <class T>
__device__ __forceinline__ T *at_offset(const void *base, T *off) {
  return reinterpret_cast<T *>(const_cast<char *>(static_cast<const char *>(base)) +
                               reinterpret_cast<intptr_t>(off));
}
// The buffers of a table entry, from the table's address (the kernel's
// table argument).
__device__ __forceinline__ Bufs bufs_at(const VolTab *vt, const Bufs &o) {
  Bufs b;
  b.u = at_offset(vt, o.u);
  b.obs = at_offset(vt, o.obs);
  b.scal = at_offset(vt, o.scal);
  b.S = at_offset(vt, o.S);
  b.K = at_offset(vt, o.K);
  b.parts = at_offset(vt, o.parts);
  b.twM = at_offset(vt, o.twM);
  b.twN = at_offset(vt, o.twN);
  b.twY = at_offset(vt, o.twY);
  b.twZ = at_offset(vt, o.twZ);
  return b;
}

// Copies a twiddle table into LDS (cooperatively; caller synchronises).
// A table of N entries moved from global memory to LDS by NT threads in two
// phases, load() into registers and store() to LDS, so that a kernel can
// issue all its loads before its first LDS store: their latencies then
// overlap instead of adding up (the vector-memory counter is in order).
template <int NT, int N>
struct TableRegs {
  static constexpr int IT = (N + NT - 1) / NT;
  float2 t[IT];
  // entry i = threadIdx.x + it * NT exists for every it but possibly the last
  __device__ __forceinline__ static bool has(int it) {
    return (it + 1) * NT <= N || (int)threadIdx.x + it * NT < N;
  }
  __device__ __forceinline__ void load(const float2 *__restrict__ src) {
#pragma unroll
    for (int it = 0; it < IT; it++)
      if (has(it)) t[it] = src[threadIdx.x + it * NT];
  }
  __device__ __forceinline__ void store(float2 *dst) const {
#pragma unroll
    for (int it = 0; it < IT; it++)
      if (has(it)) dst[threadIdx.x + it * NT] = t[it];
  }
};

// Sum of the workgroup's NT per-thread values in a fixed order: a butterfly
// across each wavefront's 64 lanes (cross-lane moves, no barrier), one LDS
// slot per wavefront, one barrier, and thread 0 adds the wavefronts' sums in
// order.  `sh` holds NT/64 doubles and must be free for the workgroup to
// overwrite (the caller synchronises before reusing it).  Returns the sum
// in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double acc, double *sh) {
  static_assert(NT % 64 == 0, "whole wavefronts");
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NT / 64; w++) s += sh[w];
  }
  return s;
}

template <int NT, int N>
__device__ __forceinline__ void stage_table(float2 *dst, const float2 *__restrict__ src) {
  TableRegs<NT, N> r;
  r.load(src);
  r.store(dst);
}

// ------------------------------------------------------------- x passes

__device__ __forceinline__ long long s_index(const Geom &g, int z, int k, int y) {
  return ((long long)z * g.hx + k) * g.ny + y;
}

// Writes X[0..M] of the TP rows held (as Z = FFT_M(x_even + i x_odd)) in
// LDS (RowLds<PS> layout, rows kRowStride<M, PS> apart), to S with consecutive
// rows on consecutive lanes.  Thread (qq, k0) = (threadIdx % TP, threadIdx /
// TP) writes row qq at k = k0 + KS*it, KS = NT/TP, unrolled: a multiple of
// the pad period, so every LDS and table address is the first one plus a
// constant.  Stores address plane z of S by 32-bit byte offsets: entries
// k > M of the last step and rows past the image have offsets at or beyond
// the plane's extent and are not stored (global stores under that
// predicate measured faster than raw buffer stores, which drop such
// entries by themselves, for the 64-byte spans of 8-row tiles).
template <int M, int TP, int NT, int PS>
__device__ __forceinline__ void unpack_store(const float2 *lds, float2 *__restrict__ S,
                                             const float2 *twN, const Geom &g, int z,
                                             int y0) {
  constexpr int MP = kRowStride<M, PS>;
  static_assert(NT % TP == 0, "a thread keeps one row");
  constexpr int KS = NT / TP;
  constexpr int IT = (M + 1 + KS - 1) / KS;
  // steps of whole pad periods (every size but 128 points at 16 per thread)
  constexpr bool WHOLE = KS % (1 << PS) == 0;
  constexpr int KSP = KS + (KS >> PS);  // then padded<PS>(k + KS) - padded<PS>(k)
  const int qq = threadIdx.x % TP;
  const int k0 = threadIdx.x / TP;
  const float2 *row = lds + qq * MP;
  const float2 *zk0 = row + padded<PS>(k0);
  const float2 *zm0 = row + padded<PS>(M - k0);
  const unsigned extent = (unsigned)g.hx * (unsigned)g.ny * (unsigned)sizeof(float2);
  char *const plane = reinterpret_cast<char *>(S + (long long)z * g.hx * g.ny);
  const unsigned off =
      (y0 + qq < g.iy) ? (unsigned)(k0 * g.ny + y0 + qq) * (unsigned)sizeof(float2) : extent;
  const unsigned stepB = (unsigned)KS * (unsigned)g.ny * (unsigned)sizeof(float2);
#pragma unroll
  for (int it = 0; it < IT; it++) {
    float2 zk, zm, w;
    if (it == 0) {
      // k = k0: Z[0]'s partner is Z[0] itself
      zk = zk0[0];
      zm = (k0 == 0) ? zk : zm0[0];
      w = twN[k0];
    } else if (it + 1 < IT) {
      if constexpr (WHOLE) {
        zk = zk0[it * KSP];
        zm = zm0[-it * KSP];
      } else {
        zk = row[padded<PS>(k0 + it * KS)];
        zm = row[padded<PS>(M - k0 - it * KS)];
      }
      w = twN[k0 + it * KS];
    } else {
      // last step: k = M (Z is M-periodic, Z[M] = Z[0], partner Z[0]) or
      // past it (dropped), read at k = M
      const int k = min(k0 + it * KS, M);
      zk = row[padded<PS>(k == M ? 0 : k)];
      zm = row[padded<PS>(M - k)];
      w = twN[k];
    }
    const float2 a = make_float2(zk.x + zm.x, zk.y - zm.y);  // Zk + conj(Zm)
    const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);  // Zk - conj(Zm)
    const float2 wd = cmul(w, d);                            // w^k d
    // X = (a - i w^k d) / 2
    const float2 X = make_float2(0.5f * (a.x + wd.y), 0.5f * (a.y - wd.x));
    const unsigned o = off + it * stepB;
    if (o < extent) *reinterpret_cast<float2 *>(plane + o) = X;
  }
}

// Loads X[0..M] of TP rows from S into LDS (row layout); the real parts of
// X[0] and X[M] share slot 0 (their imaginary parts are ignored, as a
// multi-dimensional C2R ignores them).
// Thread (qq, k0) = (threadIdx % TP, threadIdx / TP) handles row qq at
// k = k0, k0 + NT/TP, ...; load() issues all its loads, store() writes them
// to LDS (the two phases as in TableRegs).
template <int M, int TP, int NT>
struct SpectrumRegs {
  static_assert(NT % TP == 0, "a thread keeps one row");
  static constexpr int KS = NT / TP;
  static constexpr int IT = (M + 1 + KS - 1) / KS;
  float2 X[IT];
  __device__ __forceinline__ void load(const float2 *__restrict__ S, const Geom &g, int z,
                                       int y0) {
    const int qq = threadIdx.x % TP;
    const int k0 = threadIdx.x / TP;
    // unconditional loads (row and k clamped): no branches between them, so
    // all are in flight together.  No value needs masking: entries k > M are
    // never stored, and a row past the image (a partial tile) is transformed
    // on its own and never written back.
    const float2 *src = S + s_index(g, z, 0, min(y0 + qq, g.iy - 1));
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int k = k0 + it * KS;  // <= M except possibly in the last step
      X[it] = src[(long long)(it + 1 < IT ? k : min(k, M)) * g.ny];
    }
  }
  template <int PS>
  __device__ __forceinline__ void store(float2 *lds) const {
    const int qq = threadIdx.x % TP;
    const int k0 = threadIdx.x / TP;
    float2 *row = lds + qq * kRowStride<M, PS>;
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int k = k0 + it * KS;
      if (k == 0)
        row[0].x = X[it].x;
      else if (k == M)
        row[0].y = X[it].x;
      else if (k < M)
        row[padded<PS>(k)] = X[it];
    }
  }
};

// Forms Zc[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k]) for the
// thread's k = t + T*e; its inverse FFT is x[2m] + i x[2m+1].
template <int M, int E, class L_t>
__device__ __forceinline__ void pack_inverse(float2 (&v)[E], const L_t &L, int t,
                                             const float2 *twN) {
  constexpr int T = M / E;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int k = t + T * e;
    float2 xk, xm;
    if (k == 0) {
      const float2 s = L(0);
      xk = make_float2(s.x, 0.f);
      xm = make_float2(s.y, 0.f);
    } else {
      xk = L(k);
      xm = L(M - k);
    }
    const float2 a = make_float2(xk.x + xm.x, xk.y - xm.y);
    const float2 d = make_float2(xk.x - xm.x, xk.y + xm.y);
    float2 w = twN[k];
    w.y = -w.y;  // w^-k
    const float2 wd = cmul(w, d);
    v[e] = make_float2(a.x - wd.y, a.y + wd.x);  // a + i w^-k d
  }
}

// Row loads of a real image row pair (x[2m], x[2m+1]) for m = t + T*e.  With
// ix even every row starts 8-byte aligned and a pair is one 8-byte load.
// The loads are unconditional (addresses clamped into the row, rows of
// invalid tiles clamped by the caller) and the values outside the image
// zeroed by a 0/1 factor (the clamped loads read finite image values), not
// by a select, which the compiler would turn back into a branch around the
// load: without branches between them every load is in flight together.
// Element e of a thread holds the real positions n0 = 2 (t + T e), n0 + 1.
// With HE > 0 (ix == 2 T HE: a row that ends on an element boundary, e.g.
// HE = E/2 for a same-size kernel with ix = M, or HE = 2E/3 for ix = 2M/3)
// exactly the elements e < HE lie in the image, known at compile time.
template <int HE>
__device__ __forceinline__ bool in_image(int e, int n0, int ix) {
  if constexpr (HE > 0)
    return e < HE;
  else
    return n0 < ix;
}
template <int HE>
__device__ __forceinline__ bool second_in_image(int e, int n0, int ix) {
  if constexpr (HE > 0)
    return e < HE;
  else
    return n0 + 1 < ix;
}

template <int E, int T, int HE = 0>
__device__ __forceinline__ void load_row_pairs(float2 (&v)[E], const float *row, bool valid,
                                               int t, int ix) {
  if constexpr (HE > 0) {
    // rows of a partial tile are transformed alone and never stored, so
    // they need no masking; the elements past the image are zero
#pragma unroll
    for (int e = 0; e < E; e++)
      v[e] = (e < HE) ? *reinterpret_cast<const float2 *>(row + 2 * (t + T * e))
                      : make_float2(0.f, 0.f);
    (void)valid;
    (void)ix;
  } else if ((ix & 1) == 0) {
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int n0 = 2 * (t + T * e);
      const float2 p = *reinterpret_cast<const float2 *>(row + min(n0, ix - 2));
      const float m = (valid && n0 < ix) ? 1.f : 0.f;
      v[e] = make_float2(p.x * m, p.y * m);
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int n0 = 2 * (t + T * e);
      const float a = row[min(n0, ix - 1)];
      const float b = row[min(n0 + 1, ix - 1)];
      v[e] = make_float2(a * ((valid && n0 < ix) ? 1.f : 0.f),
                         b * ((valid && n0 + 1 < ix) ? 1.f : 0.f));
    }
  }
}

// Image row y of plane z for loads; rows past the image (partial tiles) are
// clamped onto the last one, whose values the loads then discard.
__device__ __forceinline__ const float *load_row(const float *base, const Geom &g, int z,
                                                 int y) {
  return base + ((long long)z * g.iy + min(y, g.iy - 1)) * g.ix;
}

// Stores (a, b) to row[n0], row[n0 + 1] (inside the row's ix entries).
__device__ __forceinline__ void store_row_pair(float *row, int n0, int ix, float a, float b) {
  if ((ix & 1) == 0) {
    *reinterpret_cast<float2 *>(row + n0) = make_float2(a, b);
  } else {
    row[n0] = a;
    if (n0 + 1 < ix) row[n0 + 1] = b;
  }
}

// x kernels share one shared-memory block: the TP-row tile, then the M-point
// and (M+1)-entry twiddle tables.
// Tile index of this workgroup.  With 16-row tiles a tile spans 128 bytes
// of every spectrum row and blocks map to tiles in order.  With 8-row tiles
// two tiles share each 128-byte span; blocks b and b + 8, which the
// dispatcher sends to the same XCD (round-robin over 8) at nearly the same
// time, take the two halves, so the line is fetched into that XCD's L2 once.
// The grid is padded to a multiple of 16; padding blocks get -1.
// With rev the tiles are taken in the opposite order (pairs of 8-row tiles
// stay pairs: tiles 2j, 2j + 1 become ntiles - 1 - 2j, ntiles - 2 - 2j).
template <int TP>
__device__ __forceinline__ int x_tile(int b, int ntiles, int rev) {
  int t;
  if constexpr (TP == 16) {
    t = b;
  } else {
    const int loc = b & 15;
    t = (b - loc) + (loc < 8 ? 2 * loc : 2 * (loc - 8) + 1);
  }
  if (t >= ntiles) return -1;
  return rev ? ntiles - 1 - t : t;
}

template <int M, int TP, int PS>
struct XSmem {
  float2 tile[TP * kRowStride<M, PS>];
  float2 twM[M];
  float2 twN[M + 1];
};

template <int M, int E, int TP, int HE>
__device__ __forceinline__ void x_fwd_body(int blk, const float *__restrict__ u, float2 *__restrict__ S,
            const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN, Geom g) {
  constexpr int T = M / E;
  constexpr int NT = TP * T;
  constexpr int PS = kPadShift<E>;
  __shared__ XSmem<M, TP, PS> sm;
  const int q = threadIdx.x / T, t = threadIdx.x % T;
  const int ntile = (g.iy + TP - 1) / TP;
  const int tile = x_tile<TP>(blk, g.iz * ntile, g.xrev);
  if (tile < 0) return;  // whole workgroup: padding block of the remap
  const int z = tile / ntile;
  const int y0 = (tile % ntile) * TP;
  const int y = y0 + q;
  TableRegs<NT, M> tm;
  TableRegs<NT, M + 1> tn;
  tm.load(gtwM);
  tn.load(gtwN);
  float2 v[E];
  load_row_pairs<E, T, HE>(v, load_row(u, g, z, y), y < g.iy, t, g.ix);
  tm.store(sm.twM);
  tn.store(sm.twN);
  __syncthreads();
  const RowLds<PS> L{sm.tile + q * kRowStride<M, PS>};
  fft_stages<M, E, 1, -1>(v, t, L, sm.twM, typename XPlan<M, E>::Rad{});
#pragma unroll
  for (int e = 0; e < E; e++) L(t + T * e) = v[e];
  __syncthreads();
  unpack_store<M, TP, NT, PS>(sm.tile, S, sm.twN, g, z, y0);
}

template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    k_x_fwd(const float *__restrict__ u, float2 *__restrict__ S,
            const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN, Geom g) {
  x_fwd_body<M, E, TP, HE>(blockIdx.x, u, S, gtwM, gtwN, g);
}
template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    kb_x_fwd(const VolTab *__restrict__ vt, const int *__restrict__ blk0s, int nvol) {
  const VolTab &v = vt[volume_of(blk0s, nvol, (int)blockIdx.x)];
  const Bufs b = bufs_at(vt, v.b);
  x_fwd_body<M, E, TP, HE>((int)blockIdx.x - v.blk0, b.u, b.S, b.twM, b.twN, v.g);
}

template <int M, int E, int TP, int HE>
__device__ __forceinline__ void x_mid_body(int blk, float2 *__restrict__ S, const float *__restrict__ obs,
            const float *__restrict__ cptr, const float2 *__restrict__ gtwM,
            const float2 *__restrict__ gtwN, Geom g) {
  constexpr int T = M / E;
  constexpr int NT = TP * T;
  constexpr int PS = kPadShift<E>;
  __shared__ XSmem<M, TP, PS> sm;
  const int q = threadIdx.x / T, t = threadIdx.x % T;
  const int ntile = (g.iy + TP - 1) / TP;
  const int tile = x_tile<TP>(blk, g.iz * ntile, g.xrev);
  if (tile < 0) return;  // whole workgroup: padding block of the remap
  const int z = tile / ntile;
  const int y0 = (tile % ntile) * TP;
  const int y = y0 + q;
  const bool valid = y < g.iy;
  // every load of the workgroup is in flight before the first LDS store;
  // the observation row is needed only after the inverse transform, so its
  // loads go last and the stores wait for the spectrum and tables only
  SpectrumRegs<M, TP, NT> sp;
  sp.load(S, g, z, y0);
  TableRegs<NT, M> tm;
  TableRegs<NT, M + 1> tn;
  tm.load(gtwM);
  tn.load(gtwN);
  float2 o[E];
  load_row_pairs<E, T, HE>(o, load_row(obs, g, z, y), valid, t, g.ix);
  sp.template store<PS>(sm.tile);
  tm.store(sm.twM);
  tn.store(sm.twN);
  __syncthreads();
  const RowLds<PS> L{sm.tile + q * kRowStride<M, PS>};
  float2 v[E];
  pack_inverse<M, E>(v, L, t, sm.twN);
  L.sync();  // the row's own lanes read it; the FFT overwrites it next
  fft_stages<M, E, 1, +1>(v, t, L, sm.twM, typename XPlan<M, E>::Rad{});
  // ratio r = guard(obs / s1) with s1 = c * (H (*) u) the blurred estimate;
  // |s1| < 1e-5 (the reference compares in double; kDivGuard is the same
  // test in float) gives 0, as the reference's ML step
  // (src/ComputePrimitives.cxx:143-148, src/MaximumLikelihoodDeconvolve.cxx:62).
  // This is guarded_ratio(o, s, 0) written out: the compiler allocates fewer
  // registers for the x pass with the expression inline.
  const float c = cptr[0];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int n0 = 2 * (t + T * e);
    float r0 = 0.f, r1 = 0.f;
    if (valid && in_image<HE>(e, n0, g.ix)) {
      const float s = c * v[e].x;
      r0 = (fabsf(s) < kDivGuard) ? 0.f : quotient(o[e].x, s);
    }
    if (valid && second_in_image<HE>(e, n0, g.ix)) {
      const float s = c * v[e].y;
      r1 = (fabsf(s) < kDivGuard) ? 0.f : quotient(o[e].y, s);
    }
    v[e] = make_float2(r0, r1);
  }
  fft_stages<M, E, 1, -1>(v, t, L, sm.twM, typename XPlan<M, E>::Rad{});
#pragma unroll
  for (int e = 0; e < E; e++) L(t + T * e) = v[e];
  __syncthreads();
  unpack_store<M, TP, NT, PS>(sm.tile, S, sm.twN, g, z, y0);
}

template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    k_x_mid(float2 *__restrict__ S, const float *__restrict__ obs,
            const float *__restrict__ cptr, const float2 *__restrict__ gtwM,
            const float2 *__restrict__ gtwN, Geom g) {
  x_mid_body<M, E, TP, HE>(blockIdx.x, S, obs, cptr, gtwM, gtwN, g);
}
template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    kb_x_mid(const VolTab *__restrict__ vt, const int *__restrict__ blk0s, int nvol) {
  const VolTab &v = vt[volume_of(blk0s, nvol, (int)blockIdx.x)];
  const Bufs b = bufs_at(vt, v.b);
  x_mid_body<M, E, TP, HE>((int)blockIdx.x - v.blk0, b.S, b.obs, b.scal + 2, b.twM, b.twN, v.g);
}

// Jansen-van Cittert update (reference src/JansenVanCittertDeconvolve.cxx:
// 106-112) with o the blurred estimate: gamma = 1 - (o - A)^2 / A^2, u' =
// max(0, u + gamma (in - o)), without contraction, as the reference's float
// code and the generic driver.
__device__ __forceinline__ float jvc_update(float u, float in, float o, float A,
                                            float invASq) {
#pragma clang fp contract(off)
  const float diff = o - A;
  const float gamma = 1.0f - ((diff * diff) * invASq);
  const float val = u + (gamma * (in - o));
  return val < 0.0f ? 0.0f : val;
}

// MODE 0: the RL update u' = (c u) s2 and its sum; MODE 1: the JVC update
// of u with the blurred estimate s2 and the observation row (cptr: A and
// 1/A^2), no sum; MODE 2: u' = s2, the pass's C2R itself (convolution and
// Wiener output).
template <int M, int E, int TP, int HE, int MODE = 0>
__device__ __forceinline__ void x_last_body(int blk, const float2 *__restrict__ S, float *__restrict__ u,
             const float *__restrict__ cptr, double *__restrict__ partials,
             const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN,
             const float *__restrict__ obs, Geom g) {
  constexpr int T = M / E;
  constexpr int NT = TP * T;
  constexpr int PS = kPadShift<E>;
  __shared__ XSmem<M, TP, PS> sm;
  const int q = threadIdx.x / T, t = threadIdx.x % T;
  const int ntile = (g.iy + TP - 1) / TP;
  const int tile = x_tile<TP>(blk, g.iz * ntile, g.xrev);
  if (tile < 0) return;  // whole workgroup: padding block of the remap
  const int z = tile / ntile;
  const int y0 = (tile % ntile) * TP;
  const int y = y0 + q;
  const bool valid = y < g.iy;
  float *urow = u + ((long long)z * g.iy + y) * g.ix;
  // loads in the order they are consumed (see k_x_mid)
  SpectrumRegs<M, TP, NT> sp;
  sp.load(S, g, z, y0);
  TableRegs<NT, M> tm;
  TableRegs<NT, M + 1> tn;
  tm.load(gtwM);
  tn.load(gtwN);
  float2 uo[E];
  load_row_pairs<E, T, HE>(uo, load_row(u, g, z, y), valid, t, g.ix);
  float2 ob[MODE == 1 ? E : 1];
  if constexpr (MODE == 1) load_row_pairs<E, T, HE>(ob, load_row(obs, g, z, y), valid, t, g.ix);
  sp.template store<PS>(sm.tile);
  tm.store(sm.twM);
  tn.store(sm.twN);
  __syncthreads();
  const RowLds<PS> L{sm.tile + q * kRowStride<M, PS>};
  float2 v[E];
  pack_inverse<M, E>(v, L, t, sm.twN);
  L.sync();  // the row's own lanes read it; the FFT overwrites it next
  fft_stages<M, E, 1, +1>(v, t, L, sm.twM, typename XPlan<M, E>::Rad{});
  if constexpr (MODE == 2) {
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int n0 = 2 * (t + T * e);
      if (valid && in_image<HE>(e, n0, g.ix))
        store_row_pair(urow, n0, g.ix, v[e].x,
                       second_in_image<HE>(e, n0, g.ix) ? v[e].y : 0.f);
    }
    return;
  }
  if constexpr (MODE == 1) {
    const float A = cptr[0], invASq = cptr[1];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int n0 = 2 * (t + T * e);
      if (valid && in_image<HE>(e, n0, g.ix)) {
        const bool second = second_in_image<HE>(e, n0, g.ix);
        const float u0 = jvc_update(uo[e].x, ob[e].x, v[e].x, A, invASq);
        const float u1 = second ? jvc_update(uo[e].y, ob[e].y, v[e].y, A, invASq) : 0.f;
        store_row_pair(urow, n0, g.ix, u0, u1);
      }
    }
    return;
  }
  // u' = (c * u) * s2: the reference's g' = g * s2 with g = c * u stored
  // rounded (src/MaximumLikelihoodDeconvolve.cxx:69-78)
  const float c = cptr[0];
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int n0 = 2 * (t + T * e);
    if (valid && in_image<HE>(e, n0, g.ix)) {
      const bool second = second_in_image<HE>(e, n0, g.ix);
      const float u0 = (c * uo[e].x) * v[e].x;
      const float u1 = second ? (c * uo[e].y) * v[e].y : 0.f;
      store_row_pair(urow, n0, g.ix, u0, u1);
      acc += (double)u0;
      if (second) acc += (double)u1;
    }
  }
  // every wavefront is done with the tile before its first slots are reused
  __syncthreads();
  const double sum = block_sum<NT>(acc, reinterpret_cast<double *>(sm.tile));
  if (threadIdx.x == 0) partials[tile] = sum;
}

template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    k_x_last(const float2 *__restrict__ S, float *__restrict__ u,
             const float *__restrict__ cptr, double *__restrict__ partials,
             const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN, Geom g) {
  x_last_body<M, E, TP, HE>(blockIdx.x, S, u, cptr, partials, gtwM, gtwN, nullptr, g);
}
template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    kb_x_last(const VolTab *__restrict__ vt, const int *__restrict__ blk0s, int nvol) {
  const VolTab &v = vt[volume_of(blk0s, nvol, (int)blockIdx.x)];
  const Bufs b = bufs_at(vt, v.b);
  x_last_body<M, E, TP, HE>((int)blockIdx.x - v.blk0, b.S, b.u, b.scal + 2, b.parts, b.twM, b.twN, nullptr, v.g);
}

// x_last of one iteration fused with x_fwd of the next: the updated rows
// u' = (c*u)*s2 are written, and while they are still in registers their R2C
// along x (the next iteration's first pass, which needs no scale: c enters
// at x_mid) goes to S.  The tile's spectrum rows are read completely before
// any of them is rewritten, and no other workgroup touches them.
template <int M, int E, int TP, int HE, int MODE = 0>
__device__ __forceinline__ void x_last_fwd_body(int blk, float2 *__restrict__ S, float *__restrict__ u,
                 const float *__restrict__ cptr, double *__restrict__ partials,
                 const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN,
             const float *__restrict__ obs, Geom g) {
  constexpr int T = M / E;
  constexpr int NT = TP * T;
  constexpr int PS = kPadShift<E>;
  __shared__ XSmem<M, TP, PS> sm;
  const int q = threadIdx.x / T, t = threadIdx.x % T;
  const int ntile = (g.iy + TP - 1) / TP;
  const int tile = x_tile<TP>(blk, g.iz * ntile, g.xrev);
  if (tile < 0) return;  // whole workgroup: padding block of the remap
  const int z = tile / ntile;
  const int y0 = (tile % ntile) * TP;
  const int y = y0 + q;
  const bool valid = y < g.iy;
  float *urow = u + ((long long)z * g.iy + y) * g.ix;
  // loads in the order they are consumed (see k_x_mid)
  SpectrumRegs<M, TP, NT> sp;
  sp.load(S, g, z, y0);
  TableRegs<NT, M> tm;
  TableRegs<NT, M + 1> tn;
  tm.load(gtwM);
  tn.load(gtwN);
  float2 uo[E];
  load_row_pairs<E, T, HE>(uo, load_row(u, g, z, y), valid, t, g.ix);
  float2 ob[MODE == 1 ? E : 1];
  if constexpr (MODE == 1) load_row_pairs<E, T, HE>(ob, load_row(obs, g, z, y), valid, t, g.ix);
  sp.template store<PS>(sm.tile);
  tm.store(sm.twM);
  tn.store(sm.twN);
  __syncthreads();
  const RowLds<PS> L{sm.tile + q * kRowStride<M, PS>};
  float2 v[E];
  pack_inverse<M, E>(v, L, t, sm.twN);
  L.sync();  // the row's own lanes read it; the FFT overwrites it next
  fft_stages<M, E, 1, +1>(v, t, L, sm.twM, typename XPlan<M, E>::Rad{});
  // the same update (and for RL the same sum) as x_last
  const float c = cptr[0];
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int n0 = 2 * (t + T * e);
    float u0 = 0.f, u1 = 0.f;
    if (valid && in_image<HE>(e, n0, g.ix)) {
      const bool second = second_in_image<HE>(e, n0, g.ix);
      if constexpr (MODE == 1) {
        u0 = jvc_update(uo[e].x, ob[e].x, v[e].x, c, cptr[1]);
        if (second) u1 = jvc_update(uo[e].y, ob[e].y, v[e].y, c, cptr[1]);
      } else {
        u0 = (c * uo[e].x) * v[e].x;
        if (second) u1 = (c * uo[e].y) * v[e].y;
      }
      store_row_pair(urow, n0, g.ix, u0, u1);
      acc += (double)u0;
      if (second) acc += (double)u1;
    }
    v[e] = make_float2(u0, u1);
  }
  fft_stages<M, E, 1, -1>(v, t, L, sm.twM, typename XPlan<M, E>::Rad{});
#pragma unroll
  for (int e = 0; e < E; e++) L(t + T * e) = v[e];
  __syncthreads();
  unpack_store<M, TP, NT, PS>(sm.tile, S, sm.twN, g, z, y0);
  if constexpr (MODE == 1) return;
  // every wavefront is done with the tile before its first slots are reused
  __syncthreads();
  const double sum = block_sum<NT>(acc, reinterpret_cast<double *>(sm.tile));
  if (threadIdx.x == 0) partials[tile] = sum;
}

template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    k_x_last_fwd(float2 *__restrict__ S, float *__restrict__ u,
                 const float *__restrict__ cptr, double *__restrict__ partials,
                 const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN, Geom g) {
  x_last_fwd_body<M, E, TP, HE>(blockIdx.x, S, u, cptr, partials, gtwM, gtwN, nullptr, g);
}
template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    kb_x_last_fwd(const VolTab *__restrict__ vt, const int *__restrict__ blk0s, int nvol) {
  const VolTab &v = vt[volume_of(blk0s, nvol, (int)blockIdx.x)];
  const Bufs b = bufs_at(vt, v.b);
  x_last_fwd_body<M, E, TP, HE>((int)blockIdx.x - v.blk0, b.S, b.u, b.scal + 2, b.parts, b.twM, b.twN, nullptr, v.g);
}

// The last x pass of a Jansen-van Cittert iteration (C2R, update) and the
// same fused with the next iteration's R2C.
template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    k_x_jvc(const float2 *__restrict__ S, float *__restrict__ u, const float *__restrict__ ap,
            const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN,
            const float *__restrict__ obs, Geom g) {
  x_last_body<M, E, TP, HE, 1>(blockIdx.x, S, u, ap, nullptr, gtwM, gtwN, obs, g);
}
template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    k_x_jvc_fwd(float2 *__restrict__ S, float *__restrict__ u, const float *__restrict__ ap,
                const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN,
                const float *__restrict__ obs, Geom g) {
  x_last_fwd_body<M, E, TP, HE, 1>(blockIdx.x, S, u, ap, nullptr, gtwM, gtwN, obs, g);
}

// The last x pass of a single convolution or Wiener filter: the C2R
// written to u as it is.
template <int M, int E, int TP, int HE>
__global__ void __launch_bounds__(TP *(M / E))
    k_x_out(const float2 *__restrict__ S, float *__restrict__ u,
            const float2 *__restrict__ gtwM, const float2 *__restrict__ gtwN, Geom g) {
  x_last_body<M, E, TP, HE, 2>(blockIdx.x, S, u, nullptr, nullptr, gtwM, gtwN, nullptr, g);
}

// ------------------------------------------------------------- y and z

// H: the image is half the work domain along y (iy = N/2, a same-size
// kernel), so the forward transform's nonzero inputs and the inverse's kept
// outputs are exactly elements e < E/2 of every thread: no bounds tests.
template <int N, int E, int PPW, int SIGN, bool H>
__device__ __forceinline__ void y_body(int blk, float2 *__restrict__ S, const float2 *__restrict__ gtw, long long npencils,
        int Lin, int Lout) {
  constexpr int T = N / E;
  constexpr int NT = PPW * T;
  // unpadded (PS beyond any position): measured faster at 1024 points, where
  // the padding costs a workgroup of occupancy per CU
  constexpr int PS = 30;
  constexpr int NP = kPadded<N, PS>;
  __shared__ float2 lds[PPW * NP];
  // only the entries the stages read: at 768 points the LDS of a workgroup
  // then allows six per CU instead of five
  constexpr int TWN = kTwNeed<N, E>;
  __shared__ float2 tw[TWN];
  const int pl = threadIdx.x / T, t = threadIdx.x % T;
  const long long p = (long long)blk * PPW + pl;
  const bool valid = p < npencils;
  float2 *base = S + p * N;
  // loads without branches (pencil and index clamped, value selected), so
  // that all of them are in flight together; a pencil past the end
  // transforms a copy of the last one and stores nothing
  const float2 *lbase = S + (valid ? p : npencils - 1) * N;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int j = t + T * e;
    if constexpr (H && SIGN < 0) {
      v[e] = (e < E / 2) ? lbase[j] : make_float2(0.f, 0.f);
    } else {
      const float2 x = lbase[min(j, Lin - 1)];
      v[e] = (j < Lin) ? x : make_float2(0.f, 0.f);
    }
  }
  stage_table<NT, TWN>(tw, gtw);
  __syncthreads();
  const PencilLds<(T <= 64), PS> L{lds + pl * NP};
  fft_stages<N, E, 1, SIGN>(v, t, L, tw, typename Plan<N, E>::Rad{});
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int j = t + T * e;
    if constexpr (H && SIGN > 0) {
      if (e < E / 2 && valid) base[j] = v[e];
    } else {
      if (valid && j < Lout) base[j] = v[e];
    }
  }
}

template <int N, int E, int PPW, int SIGN, bool H>
__global__ void __launch_bounds__(PPW *(N / E))
    k_y(float2 *__restrict__ S, const float2 *__restrict__ gtw, long long npencils,
        int Lin, int Lout) {
  y_body<N, E, PPW, SIGN, H>(blockIdx.x, S, gtw, npencils, Lin, Lout);
}
template <int N, int E, int PPW, int SIGN, bool H>
__global__ void __launch_bounds__(PPW *(N / E))
    kb_y(const VolTab *__restrict__ vt, const int *__restrict__ blk0s, int nvol) {
  // the grid walked backwards: the y passes take the volumes (and each
  // volume's pencils) in the opposite order of the x and z passes before
  // them, so they start on the data those wrote last, the likeliest to be
  // still cached
  const int blk = (int)(gridDim.x - 1 - blockIdx.x);
  const VolTab &v = vt[volume_of(blk0s, nvol, blk)];
  const Bufs b = bufs_at(vt, v.b);
  y_body<N, E, PPW, SIGN, H>(blk - v.blk0, b.S, b.twY, (long long)v.g.iz * v.g.hx, SIGN < 0 ? v.g.iy : N, SIGN < 0 ? N : v.g.iy);
}

// A workgroup transforms TP adjacent ky columns of one kx (TP * 8 bytes
// contiguous per z plane); TP lanes of a wavefront read one plane.  Each
// column of S is multiplied by the matching OTF column, the z transform of
// the kernel column of K, made here in registers (the same FFT as the
// spectrum's; the kernel, shifted by -kz/2, has its kz nonzero planes at
// z < ka and z >= nz - (kz - ka)).  HZ > 0: image and kernel have the
// same depth iz = kz = HZ * T, a whole number of elements per thread (the
// launcher uses HZ = E/2: half the work domain), so the planes of S read
// and written are exactly elements e < HZ, and the kernel's are e < HZ/2
// (planes z < kz/2) and e >= E - HZ/2 (planes z >= nz - kz/2): no bounds
// tests.
template <int N, int E, int TP, int HZ, bool WIENER = false, bool CONJ = false>
__device__ __forceinline__ void z_body(int blk, float2 *__restrict__ S, const float2 *__restrict__ K,
        const float2 *__restrict__ gtw, Geom g, float nn2 = 0.f, float eps = 0.f) {
  constexpr bool H = HZ > 0;
  static_assert(HZ % 2 == 0 && HZ <= E, "kernel planes in whole elements");
  constexpr int T = N / E;
  constexpr int NT = TP * T;
  __shared__ float2 lds[TP * N];
  __shared__ float2 tw[N];
  const int q = threadIdx.x % TP, t = threadIdx.x / TP;
  const int ntile = g.ny / TP;
  const int kx = blk / ntile;
  const int y0 = (blk % ntile) * TP;
  const long long col = (long long)kx * g.ny + y0 + q;
  // element (t + T*e) of the column: one 64-bit step of T planes per e
  const long long step = (long long)T * g.sz;
  float2 *const s0 = S + (long long)t * g.sz + col;
  const float2 *const k0 = K + (long long)t * g.sz + col;
  // Every load up front, in the order of use (the counter of outstanding
  // loads is in order): twiddles, the spectrum column, the kernel column,
  // which arrives during the spectrum's forward FFT.  The wave issues them
  // at raised priority, so that its address arithmetic is not queued
  // behind the transforms of the other workgroup on the CU (z_mod 192 ->
  // 189 us at 512x512x128).
  __builtin_amdgcn_s_setprio(3);
  TableRegs<NT, N> twr;
  twr.load(gtw);
  float2 v[E];
  {
    const float2 *p = s0;
#pragma unroll
    for (int e = 0; e < E; e++, p += step) {
      if constexpr (H)
        v[e] = (e < HZ) ? *p : make_float2(0.f, 0.f);
      else
        v[e] = (t + T * e < g.iz) ? *p : make_float2(0.f, 0.f);
    }
  }
  float2 h[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    if constexpr (H) {
      // plane z = t + T*e; upper planes sit nz - kz = (E - HZ)*T planes
      // lower in K
      if (e < HZ / 2)
        h[e] = k0[e * step];
      else if (e >= E - HZ / 2)
        h[e] = k0[(e - (E - HZ)) * step];
      else
        h[e] = make_float2(0.f, 0.f);
    } else {
      // conditional loads: planes between the two ranges take no register
      // until the FFT (unconditional loads with a 0/1 factor need ~20 more
      // VGPRs here and spill at 1024 threads)
      const int z = t + T * e;
      const int hi = g.nz - (g.kz - g.ka);  // first upper plane
      if (z < g.ka)
        h[e] = k0[e * step];
      else if (z >= hi)
        h[e] = K[(long long)(z - (g.nz - g.kz)) * g.sz + col];
      else
        h[e] = make_float2(0.f, 0.f);
    }
  }
  __builtin_amdgcn_s_setprio(0);
  twr.store(tw);
  __syncthreads();
  const InterLds<TP> L{lds + q};
  // HZ = E/2: the spectrum column is zero in its upper half, the kernel
  // column in its middle half, and the inverse is stored for its lower
  // half (patterns the first and last butterflies prune)
  constexpr bool HALF = 2 * HZ == E;
  constexpr int ZV = HALF ? kZeroUpper : 0, ZH = HALF ? kZeroMiddle : 0,
                KO = HALF ? kKeepLower : 0;
  fft_pruned<N, E, 1, -1, ZV, 0>(v, t, L, tw, typename Plan<N, E>::Rad{});
  fft_pruned<N, E, 1, -1, ZH, 0>(h, t, L, tw, typename Plan<N, E>::Rad{});
  if constexpr (WIENER) {
    // reference Wiener filter (src/WienerDeconvolve.cxx:54-64): Hc =
    // conj(H) (1 / (|H|^2 + eps)), then Hc G s.  K holds s H (s = 1/n, n
    // = nx ny nz), so Hc s = conj(s H) / (n^2 |s H|^2 + eps)
#pragma unroll
    for (int e = 0; e < E; e++) {
      const float2 hh = h[e];
      const float inv = 1.0f / (nn2 * (hh.x * hh.x + hh.y * hh.y) + eps);
      h[e] = make_float2(hh.x * inv, -hh.y * inv);
    }
  }
  // the product with the kernel column's spectrum (reference
  // ComplexMultiplyAndScale, a c - b d, a d + b c, its scale folded into
  // K), or with its conjugate for a correlation (CONJ), taken inside the
  // inverse transform's first butterflies
  fft_pruned<N, E, 1, +1, 0, KO, CONJ ? kMulConj : kMulAll>(v, t, L, tw,
                                                            typename Plan<N, E>::Rad{}, h);
  {
    float2 *p = s0;
#pragma unroll
    for (int e = 0; e < E; e++, p += step) {
      if constexpr (H) {
        if (e < HZ) *p = v[e];
      } else {
        if (t + T * e < g.iz) *p = v[e];
      }
    }
  }
}

template <int N, int E, int TP, int HZ, bool CONJ = false>
__global__ void __launch_bounds__(TP *(N / E))
    k_z(float2 *__restrict__ S, const float2 *__restrict__ K,
        const float2 *__restrict__ gtw, Geom g) {
  z_body<N, E, TP, HZ, false, CONJ>(blockIdx.x, S, K, gtw, g);
}
// batched: RL volumes only (make_rl_group takes no kernel-step engine)
template <int N, int E, int TP, int HZ>
__global__ void __launch_bounds__(TP *(N / E))
    kb_z(const VolTab *__restrict__ vt, const int *__restrict__ blk0s, int nvol) {
  const VolTab &v = vt[volume_of(blk0s, nvol, (int)blockIdx.x)];
  const Bufs b = bufs_at(vt, v.b);
  z_body<N, E, TP, HZ>((int)blockIdx.x - v.blk0, b.S, b.K, b.twZ, v.g);
}
// z pass of the Wiener filter (single volume)
template <int N, int E, int TP, int HZ>
__global__ void __launch_bounds__(TP *(N / E))
    k_zw(float2 *__restrict__ S, const float2 *__restrict__ K, const float2 *__restrict__ gtw,
         Geom g, float nn2, float eps) {
  z_body<N, E, TP, HZ, true>(blockIdx.x, S, K, gtw, g, nn2, eps);
}

// The same-size z pass (iz = kz = N/2) over several column tiles per
// workgroup, software pipelined.  Three N-point transforms per column cost
// about as much VALU time as the pass's HBM traffic, and a workgroup of the
// one-tile kernel leaves HBM idle while it computes.  Here, once the product
// is formed, the kernel column's registers are free: the next tile's
// spectrum planes (e < E/2) and kernel planes are loaded into them while
// the current tile's inverse transform runs.  K holds the kernel's planes
// z < N/4 then z >= 3N/4, so its elements e < E/2 at a column are those
// planes in order: element e < E/4 is plane t + T*e, element e >= E/4 is
// plane t + T*(e + E/2).  Tiles are dealt over the grid (tile = blockIdx.x
// + i * gridDim.x), so the workgroups resident together still read
// neighbouring 256-byte segments of each plane.
template <int N, int E, int TP>
__global__ void __launch_bounds__(TP *(N / E)) __attribute__((amdgpu_waves_per_eu(4)))
    k_zp(float2 *__restrict__ S, const float2 *__restrict__ K,
         const float2 *__restrict__ gtw, Geom g) {
  static_assert(E % 4 == 0, "kernel planes in whole elements");
  constexpr int T = N / E;
  constexpr int NT = TP * T;
  constexpr int EH = E / 2, EQ = E / 4;
  __shared__ float2 lds[TP * N];
  __shared__ float2 tw[N];
  const int q = threadIdx.x % TP, t = threadIdx.x / TP;
  const int ntile = g.ny / TP;
  const int ntot = g.hx * ntile;
  // Raw buffer accesses: a 32-bit per-lane byte offset (plane t and the
  // lane's column) plus a uniform offset per element (T*e planes), so each
  // tile's addresses take one register per lane.  The planes read and
  // written, z < N/2 of S and all N/2 of K, span under 2^32 bytes, the
  // buffers' extent (an access beyond it reads zero, not faults; the
  // launcher takes this kernel only where extent + (E/2 - 1) * stepB, the
  // largest offset sum, stays below 2^32).  The out-of-range offset of the
  // last prefetch is in the per-lane operand, so those loads fall outside
  // whether or not the range check adds the uniform one.
  const unsigned stepB = (unsigned)T * (unsigned)g.sz * (unsigned)sizeof(float2);
  const unsigned extent = (unsigned)(N / 2) * (unsigned)g.sz * (unsigned)sizeof(float2);
  const __amdgpu_buffer_rsrc_t rS = __builtin_amdgcn_make_buffer_rsrc(S, 0, (int)extent, kBufWord3);
  const __amdgpu_buffer_rsrc_t rK =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(K), 0, (int)extent, kBufWord3);
  const unsigned tB = (unsigned)t * (unsigned)g.sz * (unsigned)sizeof(float2);
  const auto offset = [&](int tile) {
    const int kx = tile / ntile, y0 = (tile - kx * ntile) * TP;
    return tB + (unsigned)(kx * g.ny + y0 + q) * (unsigned)sizeof(float2);
  };
  const auto ld = [&](__amdgpu_buffer_rsrc_t r, int e, unsigned o) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, o, e * stepB, 0));
  };
  const auto st = [&](int e, unsigned o, float2 x) {
    using B64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(rS, 0u, 0u, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B64, x), rS, o, e * stepB, 0);
  };
  TableRegs<NT, N> twr;
  twr.load(gtw);
  int tile = blockIdx.x;
  unsigned off = offset(tile);
  float2 a[EH], k[EH];  // the tile's spectrum planes and kernel planes
#pragma unroll
  for (int e = 0; e < EH; e++) a[e] = ld(rS, e, off);
#pragma unroll
  for (int e = 0; e < EH; e++) k[e] = ld(rK, e, off);
  twr.store(tw);
  __syncthreads();
  const InterLds<TP> L{lds + q};
  for (;;) {
    float2 v[E], h[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      v[e] = (e < EH) ? a[e] : make_float2(0.f, 0.f);
      h[e] = (e < EQ) ? k[e] : (e >= 3 * EQ) ? k[e - EH] : make_float2(0.f, 0.f);
    }
    fft_pruned<N, E, 1, -1, kZeroUpper, 0>(v, t, L, tw, typename Plan<N, E>::Rad{});
    fft_pruned<N, E, 1, -1, kZeroMiddle, 0>(h, t, L, tw, typename Plan<N, E>::Rad{});
    // reference ComplexMultiplyAndScale, its scale folded into K
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = cmul(v[e], h[e]);
    // Next tile's loads, here and not earlier (the scheduler would hoist
    // them above the product, with both columns still live).  After the
    // last tile the offset is the buffers' extent: the loads fall outside
    // them, return zero and fetch nothing, with no branch for the compiler
    // to sink the product behind.
    const int next = tile + (int)gridDim.x;  // the same for the whole workgroup
    const unsigned offn = next < ntot ? offset(next) : extent;
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int e = 0; e < EH; e++) a[e] = ld(rS, e, offn);
#pragma unroll
    for (int e = 0; e < EH; e++) k[e] = ld(rK, e, offn);
    fft_pruned<N, E, 1, +1, 0, kKeepLower>(v, t, L, tw, typename Plan<N, E>::Rad{});
#pragma unroll
    for (int e = 0; e < EH; e++) st(e, off, v[e]);
    if (next >= ntot) break;
    tile = next;
    off = offn;
  }
}

// ------------------------------------------------------------ dispatch

enum Kind { KX_FWD, KY_FWD, KZ_MOD, KY_INV, KX_MID, KX_LAST, KFINISH, KX_LAST_FWD, KNUM,
            // the Jansen-van Cittert engine's last x passes (timed as x_last
            // and x_last_fwd), and the output pass of a single convolution
            // or Wiener filter (timed as x_last)
            KX_JVC, KX_JVC_FWD, KX_OUT };
const char *const kNames[KNUM] = {"x_fwd", "y_fwd", "z_mod",  "y_inv",
                                  "x_mid", "x_last", "finish", "x_last_fwd"};

constexpr int kTPz = 16;  // pencils per z tile: 16 consecutive ky -> 128 B

// Launch shape of the fused kernels; defaults per size, overridable with
// CLARITY_FAST_TUNE="fe=8,ft=16,me=8,mt=8,le=8,lt=8,yp=4,ze=16,zt=16,zw=1,fx=1,xh=7"
// (elements per thread and tile rows of x_fwd, x_mid, x_last; y pencils per
// workgroup; z elements per thread, tile columns and tiles per workgroup of
// the same-size z pass; x_last_fwd chaining; same-size specialisations); any
// subset, the rest keep their defaults.
struct Tune {
  int xe[3] = {0, 0, 0};     // x_fwd, x_mid, x_last: elements per thread
  int xt[3] = {16, 16, 16};  // rows per tile
  int yp = 0, ze = 0, zt = 16;
  int zw = 1;                // same-size z pass: column tiles per workgroup
  int fx = 1;                // chain x_last into the next x_fwd
  int xr = 1;                // x passes walk their tiles from the last plane down
  int xh = 7;                // bit j: same-size specialisation of x kernel j
                             // (x_fwd, x_mid, x_last, x_last_fwd); measured
                             // slower for x_last_fwd with 8-row tiles and 8
                             // elements per thread
};

bool x_supported(int M) {
  return M == 64 || M == 96 || M == 128 || M == 192 || M == 256 || M == 384 || M == 512;
}
bool y_supported(int N) {
  return N == 128 || N == 192 || N == 256 || N == 384 || N == 512 || N == 768 ||
         N == 1024 || N == 2048;
}
bool z_supported(int N) {
  return N == 64 || N == 96 || N == 128 || N == 192 || N == 256 || N == 512;
}


// Number of x tiles (= x_last partial sums) for a tile height.
int x_tiles(const Geom &g, int tp) { return g.iz * ((g.iy + tp - 1) / tp); }

// Record mode of the launch sites, for setting up a batched iteration: a
// site reports the batched form of the kernel it would launch, with its
// grid and block size, instead of launching it.
struct LaunchRec {
  const void *fn;
  unsigned grid, block;
};
thread_local LaunchRec *g_rec = nullptr;

template <class K>
bool recorded(K *kb, unsigned grid, unsigned block) {
  if (!g_rec) return false;
  *g_rec = LaunchRec{reinterpret_cast<const void *>(kb), grid, block};
  return true;
}

template <int M, int E, int TP, int HE>
void launch_x_t(Kind kind, const Geom &g, const Bufs &b, hipStream_t s) {
  constexpr int NT = TP * (M / E);
  const int nt = x_tiles(g, TP);
  const unsigned grid = (unsigned)(TP == 16 ? nt : (nt + 15) / 16 * 16);
  if (kind == KX_FWD) {
    if (recorded(&kb_x_fwd<M, E, TP, HE>, grid, NT)) return;
    hipLaunchKernelGGL((k_x_fwd<M, E, TP, HE>), dim3(grid), dim3(NT), 0, s, b.u, b.S,
                       b.twM, b.twN, g);
  } else if (kind == KX_MID) {
    if (recorded(&kb_x_mid<M, E, TP, HE>, grid, NT)) return;
    hipLaunchKernelGGL((k_x_mid<M, E, TP, HE>), dim3(grid), dim3(NT), 0, s, b.S, b.obs,
                       b.scal + 2, b.twM, b.twN, g);
  } else if (kind == KX_LAST_FWD) {
    if (recorded(&kb_x_last_fwd<M, E, TP, HE>, grid, NT)) return;
    hipLaunchKernelGGL((k_x_last_fwd<M, E, TP, HE>), dim3(grid), dim3(NT), 0, s, b.S, b.u,
                       b.scal + 2, b.parts, b.twM, b.twN, g);
  } else if (kind == KX_LAST) {
    if (recorded(&kb_x_last<M, E, TP, HE>, grid, NT)) return;
    hipLaunchKernelGGL((k_x_last<M, E, TP, HE>), dim3(grid), dim3(NT), 0, s, b.S, b.u,
                       b.scal + 2, b.parts, b.twM, b.twN, g);
  } else if (kind == KX_JVC) {
    // A and 1/A^2 in the scalars' first two slots
    hipLaunchKernelGGL((k_x_jvc<M, E, TP, HE>), dim3(grid), dim3(NT), 0, s, b.S, b.u, b.scal,
                       b.twM, b.twN, b.obs, g);
  } else if (kind == KX_JVC_FWD) {
    hipLaunchKernelGGL((k_x_jvc_fwd<M, E, TP, HE>), dim3(grid), dim3(NT), 0, s, b.S, b.u,
                       b.scal, b.twM, b.twN, b.obs, g);
  } else if (kind == KX_OUT) {
    hipLaunchKernelGGL((k_x_out<M, E, TP, HE>), dim3(grid), dim3(NT), 0, s, b.S, b.u, b.twM,
                       b.twN, g);
  }
}

// With h, a row that ends on an element boundary of every thread (ix =
// 2 T HE) gets the kernel whose row bounds are compile-time: HE = E/2 (the
// image is half the x transform, ix == M: a same-size kernel on the
// reference domain) or HE = 2E/3 (ix == 2M/3: a same-size kernel on the
// alias-free domain of 1.5 times the image).
template <int M, int E, int TP>
void launch_x_h(int he, Kind kind, const Geom &g, const Bufs &b, hipStream_t s) {
  if (he == E / 2) return launch_x_t<M, E, TP, E / 2>(kind, g, b, s);
  if constexpr (E % 3 == 0) {
    if (he == 2 * E / 3) return launch_x_t<M, E, TP, 2 * E / 3>(kind, g, b, s);
  }
  launch_x_t<M, E, TP, 0>(kind, g, b, s);
}

template <int M, int E>
void launch_x_e(int tp, bool h, Kind kind, const Geom &g, const Bufs &b, hipStream_t s) {
  constexpr int T = M / E;
  const int he = (h && g.ix % (2 * T) == 0) ? g.ix / (2 * T) : 0;
  if (tp == 8)
    launch_x_h<M, E, 8>(he, kind, g, b, s);
  else
    launch_x_h<M, E, 16>(he, kind, g, b, s);
}

template <int M>
void launch_x(int e, int tp, bool h, Kind kind, const Geom &g, const Bufs &b,
              hipStream_t s) {
  if constexpr (M % 3 == 0) {
    launch_x_e<M, 12>(tp, h, kind, g, b, s);
  } else if constexpr (M == 64) {
    launch_x_e<M, 8>(tp, h, kind, g, b, s);
  } else {
    if (e == 16)
      launch_x_e<M, 16>(tp, h, kind, g, b, s);
    else
      launch_x_e<M, 8>(tp, h, kind, g, b, s);
  }
}

void dispatch_x(const Tune &tu, Kind kind, const Geom &g, const Bufs &b, hipStream_t s) {
  const int j = (kind == KX_FWD) ? 0 : (kind == KX_MID) ? 1
                : (kind == KX_LAST || kind == KX_JVC || kind == KX_OUT) ? 2 : 3;
  const int jt = j < 3 ? j : 2;  // x_last_fwd takes x_last's launch shape
  const int e = tu.xe[jt], tp = tu.xt[jt];
  const bool h = (tu.xh >> j) & 1;  // the launcher checks that the rows allow it
  switch (g.nx / 2) {
    case 64: launch_x<64>(e, tp, h, kind, g, b, s); break;
    case 96: launch_x<96>(e, tp, h, kind, g, b, s); break;
    case 128: launch_x<128>(e, tp, h, kind, g, b, s); break;
    case 192: launch_x<192>(e, tp, h, kind, g, b, s); break;
    case 256: launch_x<256>(e, tp, h, kind, g, b, s); break;
    case 384: launch_x<384>(e, tp, h, kind, g, b, s); break;
    case 512: launch_x<512>(e, tp, h, kind, g, b, s); break;
  }
}

template <int N, int E, int PPW>
void launch_y_p(bool fwd, const Geom &g, const Bufs &b, hipStream_t s) {
  const long long np = (long long)g.iz * g.hx;
  const unsigned grid = (unsigned)((np + PPW - 1) / PPW);
  const dim3 blk(PPW * (N / E));
  if (2 * g.iy == N) {
    if (fwd) {
      if (recorded(&kb_y<N, E, PPW, -1, true>, grid, blk.x)) return;
      hipLaunchKernelGGL((k_y<N, E, PPW, -1, true>), dim3(grid), blk, 0, s, b.S, b.twY, np,
                         g.iy, N);
    } else {
      if (recorded(&kb_y<N, E, PPW, +1, true>, grid, blk.x)) return;
      hipLaunchKernelGGL((k_y<N, E, PPW, +1, true>), dim3(grid), blk, 0, s, b.S, b.twY, np,
                         N, g.iy);
    }
  } else {
    if (fwd) {
      if (recorded(&kb_y<N, E, PPW, -1, false>, grid, blk.x)) return;
      hipLaunchKernelGGL((k_y<N, E, PPW, -1, false>), dim3(grid), blk, 0, s, b.S, b.twY, np,
                         g.iy, N);
    } else {
      if (recorded(&kb_y<N, E, PPW, +1, false>, grid, blk.x)) return;
      hipLaunchKernelGGL((k_y<N, E, PPW, +1, false>), dim3(grid), blk, 0, s, b.S, b.twY, np,
                         N, g.iy);
    }
  }
}

template <int N>
void launch_y(int ppw, bool fwd, const Geom &g, const Bufs &b, hipStream_t s) {
  constexpr int E = (N % 3 == 0) ? 12 : (N == 128 || N == 512) ? 8 : 16;
  constexpr int T = N / E;
  // pencils per workgroup: 2, 4 or 8 of them, at most 1024 threads
  if (ppw >= 8 && 8 * T <= 1024)
    launch_y_p<N, E, 8>(fwd, g, b, s);
  else if (ppw >= 4 || 2 * T < 64)
    launch_y_p<N, E, (4 * T <= 1024 ? 4 : 2)>(fwd, g, b, s);
  else
    launch_y_p<N, E, 2>(fwd, g, b, s);
}

void dispatch_y(const Tune &tu, bool fwd, const Geom &g, const Bufs &b, hipStream_t s) {
  switch (g.ny) {
    case 128: launch_y<128>(tu.yp, fwd, g, b, s); break;
    case 192: launch_y<192>(tu.yp, fwd, g, b, s); break;
    case 256: launch_y<256>(tu.yp, fwd, g, b, s); break;
    case 384: launch_y<384>(tu.yp, fwd, g, b, s); break;
    case 512: launch_y<512>(tu.yp, fwd, g, b, s); break;
    case 768: launch_y<768>(tu.yp, fwd, g, b, s); break;
    case 1024: launch_y<1024>(tu.yp, fwd, g, b, s); break;
    case 2048: launch_y<2048>(tu.yp, fwd, g, b, s); break;
  }
}

// What z_mod does with the spectrum column: the product with the OTF,
// with its conjugate (the kernel step's correlation), or the Wiener filter.
struct ZOp {
  bool conj = false;
  bool wiener = false;
  float nn2 = 0.f, eps = 0.f;  // Wiener: (nx ny nz)^2 and epsilon
};

template <int N, int E, int TP, int HZ>
void launch_z_hz(const ZOp &op, const Geom &g, const Bufs &b, unsigned grid, hipStream_t s) {
  if (op.wiener) {
    hipLaunchKernelGGL((k_zw<N, E, TP, HZ>), dim3(grid), dim3(TP * (N / E)), 0, s, b.S, b.K,
                       b.twZ, g, op.nn2, op.eps);
    return;
  }
  if (op.conj) {
    hipLaunchKernelGGL((k_z<N, E, TP, HZ, true>), dim3(grid), dim3(TP * (N / E)), 0, s, b.S,
                       b.K, b.twZ, g);
    return;
  }
  if (recorded(&kb_z<N, E, TP, HZ>, grid, TP * (N / E))) return;
  hipLaunchKernelGGL((k_z<N, E, TP, HZ>), dim3(grid), dim3(TP * (N / E)), 0, s, b.S, b.K,
                     b.twZ, g);
}

template <int N, int E, int TP>
void launch_z_t(int zw, const ZOp &op, const Geom &g, const Bufs &b, hipStream_t s) {
  const int ntot = g.hx * (g.ny / TP);
  const unsigned grid = (unsigned)ntot;
  // the pipelined kernel's 32-bit buffer offsets: extent plus N/2 - T
  // planes (its out-of-range offset plus the largest element offset)
  const unsigned long long extent = (unsigned long long)(N / 2) * g.sz * sizeof(float2);
  const bool fits32 = 2 * extent < (1ull << 32);
  const bool half = 2 * g.iz == N && 2 * g.kz == N;
  if (half && zw > 1 && fits32 && !op.conj && !op.wiener && !g_rec) {
    hipLaunchKernelGGL((k_zp<N, E, TP>), dim3((unsigned)((ntot + zw - 1) / zw)),
                       dim3(TP * (N / E)), 0, s, b.S, b.K, b.twZ, g);
    return;
  }
  if (half) return launch_z_hz<N, E, TP, E / 2>(op, g, b, grid, s);
  launch_z_hz<N, E, TP, 0>(op, g, b, grid, s);
}

// Tile widths 32 and 64 where the workgroup stays within 1024 threads and
// its tile within 128 KB of LDS, and ny is a multiple of the width.
template <int N, int E>
void launch_z_e(int tp, int zw, const ZOp &op, const Geom &g, const Bufs &b, hipStream_t s) {
  constexpr bool ok32 = 32 * (N / E) <= 1024 && 32 * N * 8 <= 131072;
  constexpr bool ok64 = 64 * (N / E) <= 1024 && 64 * N * 8 <= 131072;
  if constexpr (ok64) {
    if (tp == 64 && g.ny % 64 == 0) return launch_z_t<N, E, 64>(zw, op, g, b, s);
  }
  if constexpr (ok32) {
    if (tp >= 32 && g.ny % 32 == 0) return launch_z_t<N, E, 32>(zw, op, g, b, s);
  }
  launch_z_t<N, E, kTPz>(zw, op, g, b, s);
}

template <int N>
void launch_z(int e, int tp, int zw, const ZOp &op, const Geom &g, const Bufs &b, hipStream_t s) {
  if constexpr (N % 3 == 0) {
    launch_z_e<N, 12>(tp, zw, op, g, b, s);
  } else if constexpr (N == 64) {
    launch_z_e<N, 8>(tp, zw, op, g, b, s);
  } else {
    if (e == 8)
      launch_z_e<N, 8>(tp, zw, op, g, b, s);
    else
      launch_z_e<N, 16>(tp, zw, op, g, b, s);
  }
}

void dispatch_z(const Tune &tu, const Geom &g, const Bufs &b, hipStream_t s, const ZOp &op) {
  switch (g.nz) {
    case 64: launch_z<64>(tu.ze, tu.zt, tu.zw, op, g, b, s); break;
    case 96: launch_z<96>(tu.ze, tu.zt, tu.zw, op, g, b, s); break;
    case 128: launch_z<128>(tu.ze, tu.zt, tu.zw, op, g, b, s); break;
    case 192: launch_z<192>(tu.ze, tu.zt, tu.zw, op, g, b, s); break;
    case 256: launch_z<256>(tu.ze, tu.zt, tu.zw, op, g, b, s); break;
    case 512: launch_z<512>(tu.ze, tu.zt, tu.zw, op, g, b, s); break;
  }
}

// Default shapes, then the environment override.
Tune pick_tune(const Geom &g) {
  Tune t;
  const int xe = (g.nx / 2 == 256) ? 16 : 8;
  for (int j = 0; j < 3; j++) {
    t.xe[j] = xe;
    t.xt[j] = 16;
  }
  // 3 * 2^k rows (12 elements per thread): the general x_mid and x_last_fwd
  // need ~132 VGPRs, three 256-thread workgroups per CU with 8-row tiles,
  // one 512-thread workgroup with 16; rows that end on an element boundary
  // of every thread (ix = 2 T HE, HE = 6 or 8) need ~121, and with 16-row
  // tiles two 512-thread workgroups share a CU (measured at M = 384,
  // ix = 512: x_mid 127 us and x_last_fwd 163 us, against 160 / 190 us for
  // the general kernels with 8-row tiles)
  if ((g.nx / 2) % 3 == 0) {
    const int T = (g.nx / 2) / 12;
    const int he = g.ix % (2 * T) == 0 ? g.ix / (2 * T) : 0;
    t.xt[0] = 8;
    t.xt[1] = t.xt[2] = (he == 6 || he == 8) ? 16 : 8;
    t.xh = 15;
  }
  if (g.nx / 2 == 512) {
    t.xt[0] = t.xt[1] = t.xt[2] = 8;
    // Measured at 512x512x128: 16-row tiles read 128-byte spans of every
    // spectrum row (4.8 TB/s for the access pattern alone, 3.9 TB/s with 8
    // rows).  With 16 elements per thread the same-size x_mid and
    // x_last_fwd stay at <= 124 VGPRs, two 512-thread workgroups per CU;
    // without the specialisation they need ~140 and lose half of that.
    if (2 * g.ix == g.nx) {
      t.xe[1] = t.xe[2] = 16;
      t.xt[1] = t.xt[2] = 16;
      t.xh = 15;
    }
  }
  t.yp = 4;
  // 8 elements per thread at nz 128 only for the general z kernel; the
  // same-size one runs faster with 16 (measured at 64x256x256: 25 against
  // 32 us)
  t.ze = (g.nz == 128 && !(2 * g.iz == g.nz && 2 * g.kz == g.nz)) ? 8 : 16;
  t.zt = (g.nz >= 256) ? 32 : 16;  // measured: 256-byte plane segments at nz 256
  // at nz 512 the general z kernel (both columns' bounds tested) spills
  // with 32 columns (1024 threads, so at most 128 VGPRs); 16 do not
  if (g.nz == 512 && !(2 * g.iz == g.nz && 2 * g.kz == g.nz)) t.zt = 16;
  if (const char *s = std::getenv("CLARITY_FAST_TUNE")) {
    int v;
    const char *p = s;
    while (*p) {
      if (std::sscanf(p, "fe=%d", &v) == 1) t.xe[0] = v;
      else if (std::sscanf(p, "ft=%d", &v) == 1) t.xt[0] = v;
      else if (std::sscanf(p, "me=%d", &v) == 1) t.xe[1] = v;
      else if (std::sscanf(p, "mt=%d", &v) == 1) t.xt[1] = v;
      else if (std::sscanf(p, "le=%d", &v) == 1) t.xe[2] = v;
      else if (std::sscanf(p, "lt=%d", &v) == 1) t.xt[2] = v;
      else if (std::sscanf(p, "yp=%d", &v) == 1) t.yp = v;
      else if (std::sscanf(p, "ze=%d", &v) == 1) t.ze = v;
      else if (std::sscanf(p, "zt=%d", &v) == 1) t.zt = v;
      else if (std::sscanf(p, "zw=%d", &v) == 1) t.zw = v;
      else if (std::sscanf(p, "fx=%d", &v) == 1) t.fx = v;
      else if (std::sscanf(p, "xh=%d", &v) == 1) t.xh = v;
      else if (std::sscanf(p, "xr=%d", &v) == 1) t.xr = v;
      while (*p && *p != ',') p++;
      if (*p == ',') p++;
    }
  }
  return t;
}

// --------------------------------------------------------------- engine

// Launch shapes chosen by autotuning, per work shape, for the process
// (accessed under the library lock).
std::map<std::vector<int>, Tune> &tune_cache() {
  static std::map<std::vector<int>, Tune> cache;
  return cache;
}
std::vector<int> tune_key(const Geom &g) {
  return {g.ix, g.iy, g.iz, g.nx, g.ny, g.nz, g.kz};
}

class FastRL : public RLEngine {
 public:
  explicit FastRL(const Work &geom) {
    w_.set_geometry(geom);
    g_.ix = geom.id.x; g_.iy = geom.id.y; g_.iz = geom.id.z;
    g_.nx = geom.wd.x; g_.ny = geom.wd.y; g_.nz = geom.wd.z;
    g_.hx = g_.nx / 2 + 1;
    g_.sz = (long long)g_.hx * g_.ny;
    g_.kz = geom.kd.z;
    g_.ka = geom.kd.z - geom.kd.z / 2;
    tu_ = pick_tune(g_);
    // a shape autotuned earlier in this process keeps its choice, unless
    // the environment fixes the shapes
    auto hit = tune_cache().find(tune_key(g_));
    if (hit != tune_cache().end() && !std::getenv("CLARITY_FAST_TUNE")) tu_ = hit->second;
    nparts_ = x_tiles(g_, tu_.xt[2] == 8 ? 8 : 16);
    g_.xrev = tu_.xr;
  }
  ~FastRL() override {
    for (auto &ge : graph_exec_)
      if (ge) (void)hipGraphExecDestroy(ge);
    for (auto &ev : ev_pool_) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
  }

  // The batched form of the launch `kind` makes (z_mod: of the iteration's
  // second half when `second`), and this engine's entry of its table.
  LaunchRec batched(Kind kind) const {
    LaunchRec r{nullptr, 0, 0};
    g_rec = &r;
    switch (kind) {
      case KX_FWD:
      case KX_MID:
      case KX_LAST:
      case KX_LAST_FWD: dispatch_x(tu_, kind, g_, b_, nullptr); break;
      case KY_FWD: dispatch_y(tu_, true, g_, b_, nullptr); break;
      case KY_INV: dispatch_y(tu_, false, g_, b_, nullptr); break;
      case KZ_MOD: dispatch_z(tu_, g_, b_, nullptr, ZOp{}); break;
      default: break;
    }
    g_rec = nullptr;
    return r;
  }
  // (x passes in plane order: a batch reverses its y passes instead)
  VolTab table_entry() const {
    VolTab t{g_, b_, 0, nparts_};
    t.g.xrev = 0;
    return t;
  }

  // The kernel half-step of alternating blind RL: the correction a
  // correlation (the blurring input's conjugate spectrum), the estimate
  // renormalised to sum 1 (see make_fast_rl_kernel_step).
  void make_kernel_step() { corr_ = unit_ = true; }

  // Jansen-van Cittert instead of RL (make_fast_jvc): each iteration one
  // convolution and u' = max(0, u + gamma (obs - H (*) u)); A and 1/A^2 are
  // set after set_inputs (the first two scalar slots, unused by JVC).
  void make_jvc() { jvc_ = true; }
  // RL without the energy rescale (make_fast_rl_unscaled).
  void make_unscaled() { norescale_ = true; }
  // One convolution of the estimate (the input, by reset), or the Wiener
  // filter of it: iterate(1) writes the result to the estimate.
  void make_output(bool wiener, float eps) {
    out_ = true;
    wiener_ = wiener;
    eps_ = eps;
    const double n = (double)g_.nx * g_.ny * g_.nz;
    nn2_ = (float)(n * n);
  }
  bool is_output() const { return out_; }
  bool is_jvc() const { return jvc_; }
  bool is_unscaled() const { return norescale_; }
  bool is_kernel_step() const { return corr_; }
  ClarityResult_t set_jvc_constants(float A, float invASq, hipStream_t s) override {
    CL_TRY(k_fill(scal_.f(), 1, A, s));
    return k_fill(scal_.f() + 1, 1, invASq, s);
  }

  static bool supports(const Work &w) {
    return w.wd.x % 2 == 0 && x_supported(w.wd.x / 2) && y_supported(w.wd.y) &&
           z_supported(w.wd.z) && w.wd.y % kTPz == 0;
  }

  ClarityResult_t allocate() override {
    CL_TRY(obs_.alloc(sizeof(float) * w_.IV));
    CL_TRY(u_.alloc(sizeof(float) * w_.IV));
    CL_TRY(S_.alloc(sizeof(float2) * (size_t)g_.iz * g_.sz));
    CL_TRY(K_.alloc(sizeof(float2) * (size_t)g_.kz * g_.sz));
    // sized for the most x tiles any launch shape makes (8-row tiles)
    const int most = x_tiles(g_, 8);
    CL_TRY(parts_.alloc(sizeof(double) * (size_t)(most > kParts ? most : kParts)));
    CL_TRY(scal_.alloc(sizeof(float) * 4));
    AxisPlan pm, pn, py, pz;
    CL_TRY(axis_plan(g_.nx / 2, &pm));
    CL_TRY(axis_plan(g_.nx, &pn));
    CL_TRY(axis_plan(g_.ny, &py));
    CL_TRY(axis_plan(g_.nz, &pz));
    b_.twM = pm.tw; b_.twN = pn.tw; b_.twY = py.tw; b_.twZ = pz.tw;
    b_.u = u_.f(); b_.obs = obs_.f(); b_.scal = scal_.f();
    b_.S = S_.c(); b_.K = K_.c(); b_.parts = parts_.d();
    return CLARITY_SUCCESS;
  }

  ClarityResult_t set_inputs(const float *obs_d, const float *psf_d,
                             hipStream_t s) override {
    CL_HIP(hipMemcpyAsync(obs_.p, obs_d, sizeof(float) * w_.IV,
                          hipMemcpyDeviceToDevice, s));
    // K = the x and y transforms of the padded, -k/2 shifted kernel (the
    // OTF but for its z transform, which z_mod makes), over only the kz
    // planes the shifted kernel occupies: shifting by -kz/2 modulo kz
    // instead of nz puts work planes z < ka at p = z and the planes
    // z >= nz - kz/2 at p = z - (nz - kz), the layout z_mod reads.
    {
      // the padded kernel's buffer is kept by the engine (nx ny kz floats),
      // so set-up does not wait for the transform to free it; calls on the
      // stream reuse it in order
      const Dims pd{g_.nx, g_.ny, g_.kz};
      DevBuf &pad = pad_;
      if (!pad.p) CL_TRY(pad.alloc(sizeof(float) * (size_t)pd.voxels()));
      const int kshift[3] = {-w_.kd.x / 2, -w_.kd.y / 2, -w_.kd.z / 2};
      CL_TRY(k_pad_shift(pad.f(), pd, psf_d, w_.kd, kshift, 0.f, s));
      // the product's 1/(nx ny nz) (the reference's ComplexMultiplyAndScale
      // scale) folded into K once, so z_mod multiplies by K alone
      CL_TRY(k_scale(pad.f(), pad.f(), pd.voxels(), inv_volume(g_.nx, g_.ny, g_.nz), s));
      // the iteration's own x and y kernels over every row of the kz
      // planes (the "image" of this transform is all of them)
      Geom go = g_;
      go.ix = g_.nx;
      go.iy = g_.ny;
      go.iz = g_.kz;
      Bufs bo = b_;
      bo.u = pad.f();
      bo.S = K_.c();
      dispatch_x(tu_, KX_FWD, go, bo, s);
      dispatch_y(tu_, true, go, bo, s);
      CL_TRY(launch_check());
    }
    // E_obs: the padded observation's sum (zeros outside the corner); the
    // kernel step's target is 1 instead (the estimate renormalised to sum 1)
    CL_TRY(k_sum_partials(obs_.f(), w_.IV, parts_.d(), kParts, s));
    CL_TRY(k_finish_sum(parts_.d(), kParts, scal_.f(), nullptr, nullptr, s));
    if (unit_) CL_TRY(k_fill(scal_.f(), 1, 1.0f, s));
    return CLARITY_SUCCESS;
  }

  ClarityResult_t reset(const float *init_d, hipStream_t s) override {
    CL_HIP(hipMemcpyAsync(u_.p, init_d ? init_d : obs_.p, sizeof(float) * w_.IV,
                          hipMemcpyDeviceToDevice, s));
    CL_TRY(k_fill(scal_.f() + 2, 1, 1.0f, s));
    return CLARITY_SUCCESS;
  }

  // n RL iterations.  Chained (the default): x_fwd, then n bodies; every
  // body but the last ends in x_last_fwd, which also makes the next
  // iteration's x pass, the last in x_last.  Unchained (tune fx=0): n times
  // x_fwd + body.  Without per-kernel timing the bodies are captured once
  // into HIP graphs and replayed, which removes the host launch cost (it
  // matters for small volumes).  The graphs hold kernel nodes only; every
  // buffer they touch lives as long as the engine.
  ClarityResult_t iterate(int n, hipStream_t s) override {
    if (n <= 0) return CLARITY_SUCCESS;
    const bool chain = tu_.fx != 0;
    // a single convolution or filter is launched directly (capturing it
    // would cost more than the launches)
    if (timing_ || !use_graph_ || out_) {
      for (int it = 0; it < n; it++) {
        if (it == 0 || !chain) CL_TRY(run(KX_FWD, s));
        CL_TRY(enqueue_body(chain && it + 1 < n, s));
      }
      return CLARITY_SUCCESS;
    }
    if (!chain) {
      if (!graph_exec_[2]) CL_TRY(capture(2, &graph_exec_[2], s));
      for (int it = 0; it < n; it++) CL_HIP(hipGraphLaunch(graph_exec_[2], s));
      return CLARITY_SUCCESS;
    }
    CL_TRY(run(KX_FWD, s));
    for (int k = 0; k < 2; k++)
      if (!graph_exec_[k]) CL_TRY(capture(k, &graph_exec_[k], s));
    // runs of kGraphRun chained bodies in one graph: successive graph
    // launches leave the GPU idle ~9 us between them, the kernels inside a
    // graph follow each other without a gap
    int rest = n - 1;
    if (rest >= kGraphRun) {
      if (!graph_exec_[3]) CL_TRY(capture(3, &graph_exec_[3], s));
      for (; rest >= kGraphRun; rest -= kGraphRun) CL_HIP(hipGraphLaunch(graph_exec_[3], s));
    }
    for (; rest > 0; rest--) CL_HIP(hipGraphLaunch(graph_exec_[0], s));
    CL_HIP(hipGraphLaunch(graph_exec_[1], s));
    return CLARITY_SUCCESS;
  }

  // kind 0: chained body, 1: last body, 2: x_fwd + unchained body, 3:
  // kGraphRun chained bodies
  ClarityResult_t capture(int kind, hipGraphExec_t *out, hipStream_t s) {
    hipGraph_t graph = nullptr;
    CL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    ClarityResult_t r = (kind == 2) ? run(KX_FWD, s) : CLARITY_SUCCESS;
    for (int i = 0; i < (kind == 3 ? kGraphRun : 1) && r == CLARITY_SUCCESS; i++)
      r = enqueue_body(kind == 0 || kind == 3, s);
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (r != CLARITY_SUCCESS) {
      if (graph) (void)hipGraphDestroy(graph);
      return r;
    }
    CL_HIP(e);
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      *out = nullptr;
      return from_hip(e);
    }
    return CLARITY_SUCCESS;
  }

  // One iteration after its x_fwd: nine launches (the last x pass fused
  // with the next iteration's x_fwd when `chained`).
  ClarityResult_t enqueue_body(bool chained, hipStream_t s) {
    if (out_) {
      // one convolution (or Wiener filter), its C2R the result
      CL_TRY(run(KY_FWD, s));
      CL_TRY(run(KZ_MOD, s));
      CL_TRY(run(KY_INV, s));
      return run(KX_OUT, s);
    }
    if (jvc_) {
      // one convolution, then the JVC update in the last x pass
      CL_TRY(run(KY_FWD, s));
      CL_TRY(run(KZ_MOD, s));
      CL_TRY(run(KY_INV, s));
      return run(chained ? KX_JVC_FWD : KX_JVC, s);
    }
    for (int half = 0; half < 2; half++) {
      CL_TRY(run(KY_FWD, s));
      corr_now_ = corr_ && half == 1;
      const ClarityResult_t r = run(KZ_MOD, s);
      corr_now_ = false;
      CL_TRY(r);
      CL_TRY(run(KY_INV, s));
      CL_TRY(run(half == 0 ? KX_MID : chained ? KX_LAST_FWD : KX_LAST, s));
    }
    // without rescale the scale c stays at reset's 1: no finish
    return norescale_ ? CLARITY_SUCCESS : run(KFINISH, s);
  }

  ClarityResult_t estimate(float *out_d, hipStream_t s) override {
    return k_scale_dev(out_d, u_.f(), w_.IV, scal_.f() + 2, s);
  }

  const float *scale_ptr() const override { return scal_.f() + 2; }
  const char *name() const override { return "fast"; }

  int describe(const char **names, double *bytes, int *launches,
               int cap) override {
    const double f = 4.0, c = 8.0;
    const double IV = (double)w_.IV;
    const double Sp = (double)g_.iz * g_.hx * g_.iy;  // spectrum, y < iy
    const double Sf = (double)g_.iz * g_.hx * g_.ny;  // spectrum, all ky
    const double Kc = (double)g_.kz * g_.hx * g_.ny;  // kernel planes of K
    const double b[KNUM] = {
        f * IV + c * Sp,                // x_fwd
        c * Sp + c * Sf,                // y_fwd
        c * Sf + c * Kc + c * Sf,       // z_mod
        c * Sf + c * Sp,                // y_inv
        c * Sp + f * IV + c * Sp,       // x_mid
        c * Sp + f * IV + f * IV,       // x_last
        8.0 * nparts_ + 12.0,           // finish
        2 * c * Sp + 2 * f * IV,        // x_last_fwd
    };
    // launches per iteration of a long run: x_fwd and x_last run once per
    // iterate call, x_last_fwd joins the iterations in between
    const int per_it[KNUM] = {0, 2, 2, 2, 1, 0, 1, 1};
    int n = 0;
    for (int k = 0; k < KNUM && n < cap; k++, n++) {
      names[n] = kNames[k];
      bytes[n] = b[k];
      launches[n] = per_it[k];
    }
    return n;
  }

  ClarityResult_t set_timing(bool on) override {
    timing_ = on;
    return CLARITY_SUCCESS;
  }

  // Coordinate descent over the launch-shape knobs, each candidate timed
  // with HIP events over `iterations` iterations (after one warm-up
  // iteration) on the plan's own inputs; a candidate replaces the current
  // shape only when it is at least 1 % faster.  The estimate and its scale
  // are saved first and restored after, so tuning leaves no trace in the
  // results.  The choice is cached per work shape for the process and used
  // by later plans of that shape.
  ClarityResult_t autotune(int iterations, char *desc, int cap, hipStream_t s) override {
    if (iterations < 1) return CLARITY_INVALID_ARGUMENT;
    DevBuf su, ss;
    CL_TRY(su.alloc(sizeof(float) * w_.IV));
    CL_TRY(ss.alloc(sizeof(float) * 4));
    CL_HIP(hipMemcpyAsync(su.p, u_.p, sizeof(float) * w_.IV, hipMemcpyDeviceToDevice, s));
    CL_HIP(hipMemcpyAsync(ss.p, scal_.p, sizeof(float) * 4, hipMemcpyDeviceToDevice, s));
    const bool timing = timing_;
    timing_ = false;
    const bool graph = use_graph_;
    use_graph_ = false;  // time plain launches; graphs are rebuilt after
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ClarityResult_t r = from_hip(hipEventCreate(&e0));
    if (r == CLARITY_SUCCESS) r = from_hip(hipEventCreate(&e1));
    if (r == CLARITY_SUCCESS) r = tune_search(iterations, e0, e1, s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    use_graph_ = graph;
    timing_ = timing;
    CL_HIP(hipMemcpyAsync(u_.p, su.p, sizeof(float) * w_.IV, hipMemcpyDeviceToDevice, s));
    CL_HIP(hipMemcpyAsync(scal_.p, ss.p, sizeof(float) * 4, hipMemcpyDeviceToDevice, s));
    CL_HIP(hipStreamSynchronize(s));
    if (desc && cap > 0)
      std::snprintf(desc, (size_t)cap,
                    "fe=%d,ft=%d,me=%d,mt=%d,le=%d,lt=%d,yp=%d,ze=%d,zt=%d,zw=%d,fx=%d,xh=%d",
                    tu_.xe[0], tu_.xt[0], tu_.xe[1], tu_.xt[1], tu_.xe[2], tu_.xt[2], tu_.yp,
                    tu_.ze, tu_.zt, tu_.zw, tu_.fx, tu_.xh);
    return r;
  }

  int timings(int *count, double *ms, int cap, bool reset) override {
    Runtime &r = rt();
    if (r.stream) (void)hipStreamSynchronize(r.stream);
    for (auto &rec : pending_) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, ev_pool_[rec.second].first,
                              ev_pool_[rec.second].second) == hipSuccess) {
        cnt_[rec.first]++;
        ms_[rec.first] += t;
      }
    }
    pending_.clear();
    ev_next_ = 0;
    int n = 0;
    for (int k = 0; k < KNUM && n < cap; k++, n++) {
      count[n] = cnt_[k];
      ms[n] = ms_[k];
    }
    if (reset) {
      for (int k = 0; k < KNUM; k++) {
        cnt_[k] = 0;
        ms_[k] = 0.0;
      }
    }
    return n;
  }

 private:
  // Makes `t` the launch shape: captured graphs describe the old one.
  void apply_tune(const Tune &t) {
    tu_ = t;
    nparts_ = x_tiles(g_, tu_.xt[2] == 8 ? 8 : 16);
    g_.xrev = tu_.xr;
    for (auto &ge : graph_exec_) {
      if (ge) (void)hipGraphExecDestroy(ge);
      ge = nullptr;
    }
  }

  // Milliseconds per iteration of shape `t` (estimate reset to the
  // observation, one warm-up iteration, then `iterations` timed).
  ClarityResult_t time_tune(const Tune &t, int iterations, hipEvent_t e0, hipEvent_t e1,
                            hipStream_t s, float *ms) {
    apply_tune(t);
    CL_TRY(reset(nullptr, s));
    CL_TRY(iterate(1, s));
    CL_HIP(hipEventRecord(e0, s));
    CL_TRY(iterate(iterations, s));
    CL_HIP(hipEventRecord(e1, s));
    CL_HIP(hipEventSynchronize(e1));
    CL_HIP(hipEventElapsedTime(ms, e0, e1));
    *ms /= (float)iterations;
    return CLARITY_SUCCESS;
  }

  ClarityResult_t tune_search(int iterations, hipEvent_t e0, hipEvent_t e1, hipStream_t s) {
    auto &cache = tune_cache();
    const std::vector<int> key = tune_key(g_);
    auto hit = cache.find(key);
    if (hit != cache.end()) {
      apply_tune(hit->second);
      return CLARITY_SUCCESS;
    }
    Tune best = tu_;
    float best_ms = 0.f;
    CL_TRY(time_tune(best, iterations, e0, e1, s, &best_ms));
    // Each knob and the values it may take; invalid combinations fall back
    // to valid launch shapes inside the dispatchers.  The x knobs set
    // elements per thread and tile rows together (value e * 100 + rows), the
    // x_last knob also the same-size flag of x_last_fwd (+ 1): these only
    // pay off jointly (the register count decides the occupancy).
    struct Knob {
      void (*set)(Tune &, int);
      std::vector<int> values;
    };
    const std::vector<Knob> knobs = {
        {[](Tune &t, int v) { t.xe[0] = v / 100, t.xt[0] = v % 100; }, {808, 816, 1608, 1616}},
        {[](Tune &t, int v) { t.xe[1] = v / 100, t.xt[1] = v % 100; }, {808, 816, 1608, 1616}},
        {[](Tune &t, int v) {
           t.xe[2] = v / 1000, t.xt[2] = (v / 10) % 100, t.xh = (t.xh & 7) | ((v % 10) << 3);
         },
         {8080, 8081, 8160, 8161, 16080, 16081, 16160, 16161}},
        {[](Tune &t, int v) { t.xh = (t.xh & 8) | v; }, {0, 7}},
        {[](Tune &t, int v) { t.yp = v; }, {2, 4, 8}},
        {[](Tune &t, int v) { t.ze = v; }, {8, 16}},
        {[](Tune &t, int v) { t.zt = v; }, {16, 32, 64}},
        {[](Tune &t, int v) { t.zw = v; }, {1, 2, 4, 8}},
        {[](Tune &t, int v) { t.fx = v; }, {0, 1}},
    };
    auto same = [](const Tune &a, const Tune &b) {
      return std::memcmp(&a, &b, sizeof(Tune)) == 0;
    };
    // coordinate descent, a second sweep when the first moved
    for (int sweep = 0; sweep < 2; sweep++) {
      const Tune start = best;
      for (const Knob &k : knobs) {
        const Tune base = best;
        for (int v : k.values) {
          Tune c = base;
          k.set(c, v);
          if (same(c, base)) continue;
          float ms = 0.f;
          CL_TRY(time_tune(c, iterations, e0, e1, s, &ms));
          if (ms < 0.99f * best_ms) {
            best = c;
            best_ms = ms;
          }
        }
      }
      if (same(best, start)) break;
    }
    apply_tune(best);
    cache[key] = best;
    return CLARITY_SUCCESS;
  }

  ClarityResult_t run(Kind kind, hipStream_t s) {
    int slot = -1;
    if (timing_) {
      if (ev_next_ == ev_pool_.size()) {
        hipEvent_t a, b;
        CL_HIP(hipEventCreate(&a));
        CL_HIP(hipEventCreate(&b));
        ev_pool_.push_back({a, b});
      }
      slot = (int)ev_next_++;
      CL_HIP(hipEventRecord(ev_pool_[slot].first, s));
    }
    switch (kind) {
      case KX_FWD:
      case KX_MID:
      case KX_LAST:
      case KX_LAST_FWD:
      case KX_JVC:
      case KX_JVC_FWD:
      case KX_OUT: dispatch_x(tu_, kind, g_, b_, s); break;
      case KY_FWD: dispatch_y(tu_, true, g_, b_, s); break;
      case KY_INV: dispatch_y(tu_, false, g_, b_, s); break;
      case KZ_MOD: {
        ZOp op;
        op.conj = corr_now_;
        op.wiener = wiener_;
        op.nn2 = nn2_;
        op.eps = eps_;
        dispatch_z(tu_, g_, b_, s, op);
        break;
      }
      case KFINISH:
        CL_TRY(k_finish_sum(b_.parts, nparts_, b_.scal + 1, b_.scal, b_.scal + 2, s));
        break;
      default: break;
    }
    CL_TRY(launch_check());
    if (timing_) {
      CL_HIP(hipEventRecord(ev_pool_[slot].second, s));
      pending_.push_back({(kind == KX_JVC || kind == KX_OUT) ? KX_LAST
                          : kind == KX_JVC_FWD ? KX_LAST_FWD : kind,
                          slot});
    }
    return CLARITY_SUCCESS;
  }

  Geom g_{};
  Tune tu_{};
  bool use_graph_ = std::getenv("CLARITY_NO_GRAPH") == nullptr;
  hipGraphExec_t graph_exec_[4] = {nullptr, nullptr, nullptr, nullptr};  // see capture()
  static constexpr int kGraphRun = 8;  // chained bodies per graph of kind 3
  Bufs b_{};
  int nparts_ = 0;
  DevBuf obs_, u_, S_, K_, parts_, scal_, pad_;
  bool corr_ = false;      // kernel step: the correction correlates
  bool unit_ = false;      // kernel step: renormalise to sum 1
  bool corr_now_ = false;  // the z_mod being enqueued takes the conjugate
  bool jvc_ = false;       // Jansen-van Cittert iterations
  bool norescale_ = false; // RL without the energy rescale
  bool out_ = false;       // one convolution / Wiener filter
  bool wiener_ = false;
  float nn2_ = 0.f, eps_ = 0.f;
  bool timing_ = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool_;
  size_t ev_next_ = 0;
  std::vector<std::pair<int, int>> pending_;
  int cnt_[KNUM] = {};
  double ms_[KNUM] = {};
};

// The finish of several volumes, one block each: the sum of the volume's
// partials in the order and rounding of the single-volume finish (256
// threads, strided then a tree), S = (float)sum, c = E / S.
__global__ void __launch_bounds__(256) kb_finish(const VolTab *__restrict__ vt) {
  __shared__ double sh[256];
  const VolTab &v = vt[blockIdx.x];
  const Bufs b = bufs_at(vt, v.b);
  double acc = 0.0;
  for (int i = threadIdx.x; i < v.nparts; i += blockDim.x) acc += b.parts[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float f = (float)sh[0];
    b.scal[1] = f;
    b.scal[2] = b.scal[0] / f;
  }
}

// RL iterations of several fused engines together: for every pass, one
// launch per kernel instantiation over all the volumes that use it (a
// block-to-volume map and a table of the volumes' geometry and buffers),
// instead of one launch per volume; every volume's finish in one launch.
// Each volume runs the very kernels, arguments and order of its own
// engine's chained iteration, so the results are those of its engine.
class FastGroup : public RLGroup {
 public:
  explicit FastGroup(std::vector<FastRL *> e) : e_(std::move(e)) {}
  ~FastGroup() override {
    for (auto &ge : graph_)
      if (ge) (void)hipGraphExecDestroy(ge);
  }

  ClarityResult_t build() {
    const Kind kinds[] = {KX_FWD, KY_FWD, KZ_MOD, KY_INV, KX_MID, KX_LAST, KX_LAST_FWD};
    for (Kind k : kinds)
      for (int half = 0; half < 2; half++) {
        if (half == 1 && k != KZ_MOD) continue;
        // volumes by (batched kernel, block size), in volume order
        std::vector<std::pair<LaunchRec, std::vector<int>>> groups;
        for (int v = 0; v < (int)e_.size(); v++) {
          const LaunchRec r = e_[v]->batched(k);
          if (!r.fn || r.grid == 0) return CLARITY_INVALID_OPERATION;
          auto it = groups.begin();
          for (; it != groups.end(); ++it)
            if (it->first.fn == r.fn && it->first.block == r.block) break;
          if (it == groups.end()) {
            groups.push_back({r, {}});
            it = groups.end() - 1;
          }
          it->second.push_back(v);
        }
        for (auto &gr : groups) {
          auto L = std::unique_ptr<Launch>(new Launch);
          L->fn = gr.first.fn;
          L->block = gr.first.block;
          std::vector<VolTab> tab;
          std::vector<int> blk0s;
          unsigned grid = 0;
          for (int v : gr.second) {
            VolTab t = e_[v]->table_entry();
            t.blk0 = (int)grid;
            blk0s.push_back(t.blk0);
            grid += e_[v]->batched(k).grid;
            tab.push_back(t);
          }
          L->grid = grid;
          L->nvol = (int)tab.size();
          CL_TRY(L->tab.alloc(sizeof(VolTab) * tab.size()));
          CL_TRY(L->blk0s.alloc(sizeof(int) * blk0s.size()));
          for (VolTab &t : tab) t = relative_to(t, L->tab.p);
          CL_HIP(hipMemcpy(L->tab.p, tab.data(), sizeof(VolTab) * tab.size(),
                           hipMemcpyHostToDevice));
          CL_HIP(hipMemcpy(L->blk0s.p, blk0s.data(), sizeof(int) * blk0s.size(),
                           hipMemcpyHostToDevice));
          launches_[k][half].push_back(std::move(L));
        }
      }
    std::vector<VolTab> fin;
    for (FastRL *e : e_) fin.push_back(e->table_entry());
    CL_TRY(fin_.alloc(sizeof(VolTab) * fin.size()));
    for (VolTab &t : fin) t = relative_to(t, fin_.p);
    CL_HIP(hipMemcpy(fin_.p, fin.data(), sizeof(VolTab) * fin.size(), hipMemcpyHostToDevice));
    return CLARITY_SUCCESS;
  }

  // n chained iterations of every volume, as FastRL::iterate: x_fwd, then
  // n bodies, the last ending in x_last (replayed from two captured graphs
  // unless CLARITY_NO_GRAPH is set).
  ClarityResult_t iterate(int n, hipStream_t s) override {
    if (n <= 0) return CLARITY_SUCCESS;
    CL_TRY(pass(KX_FWD, 0, s));
    if (std::getenv("CLARITY_NO_GRAPH")) {
      for (int it = 0; it < n; it++) CL_TRY(body(it + 1 < n, s));
      return CLARITY_SUCCESS;
    }
    for (int k = 0; k < 2; k++)
      if (!graph_[k]) CL_TRY(capture(k == 0, &graph_[k], s));
    // runs of kGraphRun chained bodies per graph (no idle gap between the
    // graph's kernels; ~9 us between successive graphs)
    int rest = n - 1;
    if (rest >= kGraphRun) {
      if (!graph_[2]) CL_TRY(capture(true, &graph_[2], s, kGraphRun));
      for (; rest >= kGraphRun; rest -= kGraphRun) CL_HIP(hipGraphLaunch(graph_[2], s));
    }
    for (; rest > 0; rest--) CL_HIP(hipGraphLaunch(graph_[0], s));
    CL_HIP(hipGraphLaunch(graph_[1], s));
    return CLARITY_SUCCESS;
  }

 private:
  struct Launch {
    const void *fn = nullptr;
    unsigned grid = 0, block = 0;
    int nvol = 0;
    DevBuf tab, blk0s;  // volumes' entries; their first blocks, ascending
  };

  ClarityResult_t pass(Kind k, int half, hipStream_t s) {
    for (auto &L : launches_[k][half]) {
      const VolTab *t = static_cast<const VolTab *>(L->tab.p);
      const int *b = static_cast<const int *>(L->blk0s.p);
      int n = L->nvol;
      void *args[] = {&t, &b, &n};
      CL_HIP(hipLaunchKernel(L->fn, dim3(L->grid), dim3(L->block), args, 0, s));
    }
    return CLARITY_SUCCESS;
  }

  ClarityResult_t body(bool chained, hipStream_t s) {
    for (int half = 0; half < 2; half++) {
      CL_TRY(pass(KY_FWD, 0, s));
      CL_TRY(pass(KZ_MOD, half, s));
      CL_TRY(pass(KY_INV, 0, s));
      CL_TRY(pass(half == 0 ? KX_MID : chained ? KX_LAST_FWD : KX_LAST, 0, s));
    }
    const VolTab *t = static_cast<const VolTab *>(fin_.p);
    hipLaunchKernelGGL(kb_finish, dim3((unsigned)e_.size()), dim3(256), 0, s, t);
    return launch_check();
  }

  ClarityResult_t capture(bool chained, hipGraphExec_t *out, hipStream_t s, int bodies = 1) {
    hipGraph_t graph = nullptr;
    CL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    ClarityResult_t r = CLARITY_SUCCESS;
    for (int i = 0; i < bodies && r == CLARITY_SUCCESS; i++) r = body(chained, s);
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (r != CLARITY_SUCCESS) {
      if (graph) (void)hipGraphDestroy(graph);
      return r;
    }
    CL_HIP(e);
    const hipError_t ei = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      *out = nullptr;
      return from_hip(ei);
    }
    return CLARITY_SUCCESS;
  }

  std::vector<FastRL *> e_;
  std::vector<std::unique_ptr<Launch>> launches_[KNUM][2];
  DevBuf fin_;
  hipGraphExec_t graph_[3] = {nullptr, nullptr, nullptr};  // chained, last, kGraphRun chained
  static constexpr int kGraphRun = 8;
};

}  // namespace

// Work domain of the fused engine for an RL plan.
//
// Along each axis the reference transforms over n = i + k (image plus
// kernel, src/MaximumLikelihoodDeconvolve.cxx:271-289).  Every quantity an
// RL iteration keeps is zero outside the image corner [0, i) (the estimate
// starts as the zero-padded observation and every update multiplies it; the
// ratio is obs/s1 with obs = 0 there) and s1, s2 are read only inside it.
// So each circular convolution is only evaluated at offsets d = p - q with
// p, q in [0, i), |d| < i, and it equals the reference's wherever the
// kernel, shifted by -k/2 to occupy offsets [-(k/2), k - k/2), cannot reach
// such a d through the wrap-around: for a transform length N that holds iff
// N >= i + k/2 (positive offsets up to i - 1 must stay below N - k/2, and
// negative ones down to -(i - 1) above k - k/2 - 1 - N).  The reference's
// n = i + k satisfies it; so does any such N, with the same RL iterates in
// the image corner (up to the float rounding of a different FFT length).
// The engine also keeps N >= k, so that the kz kernel planes it stores
// never overlap, and takes the smallest such N it has kernels for: at a same-size
// kernel 1.5 i instead of 2 i per axis (768x768x192 instead of
// 1024x1024x256 at 512x512x128, 2.4x fewer work voxels), and shapes whose
// i + k is no size of the engine's still run on it.  CLARITY_REFERENCE_DOMAIN
// keeps n = i + k (when the engine has that size).
namespace {

int smallest_size(long long need, const int *sizes, int count) {
  for (int j = 0; j < count; j++)
    if (sizes[j] >= need) return sizes[j];
  return 0;
}

bool rl_domain(const Work &ref, Dims *wd) {
  if (std::getenv("CLARITY_REFERENCE_DOMAIN")) {
    *wd = ref.wd;
    return true;
  }
  static const int xs[] = {128, 192, 256, 384, 512, 768, 1024};
  static const int ys[] = {128, 192, 256, 384, 512, 768, 1024, 2048};
  static const int zs[] = {64, 96, 128, 192, 256, 512};
  auto need = [](int i, int k) {
    const long long a = (long long)i + k / 2;
    return a > k ? a : (long long)k;
  };
  wd->x = smallest_size(need(ref.id.x, ref.kd.x), xs, 7);
  wd->y = smallest_size(need(ref.id.y, ref.kd.y), ys, 8);
  wd->z = smallest_size(need(ref.id.z, ref.kd.z), zs, 6);
  // z_mod reads and writes only the iz image and kz kernel planes whatever
  // nz is, so a longer z transform costs FFT arithmetic, not bytes: take
  // nz = 2 iz when image and kernel are both that half (the same-size
  // specialisation of the z kernel)
  const int h = 2 * ref.id.z;
  if (ref.kd.z == ref.id.z && h >= wd->z && smallest_size(h, zs, 6) == h) wd->z = h;
  return wd->x && wd->y && wd->z;
}

}  // namespace

std::unique_ptr<RLEngine> make_fast_rl(const Work &ref) {
  Work geom;
  geom.set_geometry(ref);
  if (!rl_domain(ref, &geom.wd)) return nullptr;
  geom.V = geom.wd.voxels();
  geom.C = (long long)geom.wd.z * geom.wd.y * (geom.wd.x / 2 + 1);
  geom.scale = inv_volume(geom.wd.x, geom.wd.y, geom.wd.z);
  if (!FastRL::supports(geom)) return nullptr;
  return std::unique_ptr<RLEngine>(new FastRL(geom));
}

// The kernel half-step of alternating blind RL, h'(z) = h(z) sum_y r(y)
// O(y - z) with r = guard(obs / (h (*) O)), then h renormalised to sum 1,
// as an RL iteration of this engine: the estimate is the kernel in its
// caller layout (kernelDim, in the corner), the blurring "kernel" input is
// the image estimate O (imageDim).  Holding h in the corner instead of
// shifted by -k/2 makes the first convolution the engine's own (O shifted
// by -i/2, the same shift when i = k); the correction is the correlation
// of r with that shifted O, i.e. the second product takes conj of its
// spectrum, which z_mod applies.  On the reference work domain (n = i + k
// = 2i on every axis) the correlation of r (zero outside the image corner)
// with O does not wrap around, so the result is the generic driver's.
// nullptr unless image and kernel have the same even dimensions and the
// domain is supported.
std::unique_ptr<RLEngine> make_fast_rl_kernel_step(const Work &ref) {
  const Dims &i = ref.id, &k = ref.kd;
  if (i.x != k.x || i.y != k.y || i.z != k.z) return nullptr;
  if (k.x % 2 || k.y % 2 || k.z % 2) return nullptr;
  if (ref.wd.x != 2 * k.x || ref.wd.y != 2 * k.y || ref.wd.z != 2 * k.z) return nullptr;
  Work geom;
  geom.set_geometry(ref);
  geom.wd = ref.wd;
  geom.V = geom.wd.voxels();
  geom.C = (long long)geom.wd.z * geom.wd.y * (geom.wd.x / 2 + 1);
  geom.scale = inv_volume(geom.wd.x, geom.wd.y, geom.wd.z);
  if (!FastRL::supports(geom)) return nullptr;
  auto *e = new FastRL(geom);
  e->make_kernel_step();
  return std::unique_ptr<RLEngine>(e);
}

// Jansen-van Cittert on the fused passes: its estimate spreads over the
// whole work domain (u + gamma (obs - H (*) u) is not zero outside the
// observation), so the engine's "image" is the reference domain itself
// (every row, pencil and plane transformed) and there is no smaller
// alias-free domain; the kernel planes and K are as for RL.  nullptr when
// the domain is not one of the engine's sizes.
std::unique_ptr<RLEngine> make_fast_jvc(const Work &ref) {
  Work geom;
  geom.set_geometry(ref);
  geom.id = ref.wd;
  geom.wd = ref.wd;
  geom.IV = geom.wd.voxels();
  geom.V = geom.wd.voxels();
  geom.C = (long long)geom.wd.z * geom.wd.y * (geom.wd.x / 2 + 1);
  geom.scale = inv_volume(geom.wd.x, geom.wd.y, geom.wd.z);
  if (!FastRL::supports(geom)) return nullptr;
  auto *e = new FastRL(geom);
  e->make_jvc();
  return std::unique_ptr<RLEngine>(e);
}

// RL without the energy rescale on the fused engine (its own alias-free
// domain), for the as-written blind ML, whose image update has none.
std::unique_ptr<RLEngine> make_fast_rl_unscaled(const Work &ref) {
  std::unique_ptr<RLEngine> e = make_fast_rl(ref);
  if (e) static_cast<FastRL *>(e.get())->make_unscaled();
  return e;
}

// A single convolution on the fused passes: RL's first convolution, so
// over the same alias-free domain (the output is the image corner).
std::unique_ptr<RLEngine> make_fast_convolve(const Work &ref) {
  std::unique_ptr<RLEngine> e = make_fast_rl(ref);
  if (e) static_cast<FastRL *>(e.get())->make_output(false, 0.f);
  return e;
}

// The Wiener filter on the fused passes, over the reference domain: the
// filter conj(H) / (|H|^2 + eps) has no compact support, so no smaller
// domain gives the reference's result.
std::unique_ptr<RLEngine> make_fast_wiener(const Work &ref, float eps) {
  Work geom;
  geom.set_geometry(ref);
  geom.wd = ref.wd;
  geom.V = geom.wd.voxels();
  geom.C = (long long)geom.wd.z * geom.wd.y * (geom.wd.x / 2 + 1);
  geom.scale = inv_volume(geom.wd.x, geom.wd.y, geom.wd.z);
  if (!FastRL::supports(geom)) return nullptr;
  auto *e = new FastRL(geom);
  e->make_output(true, eps);
  return std::unique_ptr<RLEngine>(e);
}

std::unique_ptr<RLGroup> make_rl_group(const std::vector<RLEngine *> &engines) {
  std::vector<FastRL *> f;
  for (RLEngine *e : engines) {
    auto *x = dynamic_cast<FastRL *>(e);
    if (!x || x->is_jvc() || x->is_unscaled() || x->is_output() || x->is_kernel_step())
      return nullptr;
    f.push_back(x);
  }
  if (f.empty() || f.size() > 65535) return nullptr;
  auto g = std::unique_ptr<FastGroup>(new FastGroup(std::move(f)));
  if (g->build() != CLARITY_SUCCESS) return nullptr;
  return g;
}

}  // namespace clarity
