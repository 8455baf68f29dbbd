// Complex helpers and small DFT butterflies shared by the FFT kernels.
// sign = -1 forward (w = exp(-2 pi i / R)), +1 inverse; no normalisation.
#pragma once

#include <hip/hip_runtime.h>

namespace clarity {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (sign * i)
__device__ __forceinline__ float2 cmul_si(float2 a, int sign) {
  return sign < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

__device__ __forceinline__ void dft2(float2 &a, float2 &b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <typename Arr>
__device__ __forceinline__ void dft4(Arr &u, int sign) {
  float2 t0 = cadd(u[0], u[2]);
  float2 t1 = csub(u[0], u[2]);
  float2 t2 = cadd(u[1], u[3]);
  float2 t3 = cmul_si(csub(u[1], u[3]), sign);
  u[0] = cadd(t0, t2);
  u[2] = csub(t0, t2);
  u[1] = cadd(t1, t3);
  u[3] = csub(t1, t3);
}

// 8-point DFT as two 4-point DFTs (even / odd inputs) and a radix-2 merge
// with w8^k = exp(sign * i * pi * k / 4).
template <typename Arr>
__device__ __forceinline__ void dft8(Arr &u, int sign) {
  constexpr float c = 0.70710678118654752440f;
  float2 e[4] = {u[0], u[2], u[4], u[6]};
  float2 o[4] = {u[1], u[3], u[5], u[7]};
  dft4(e, sign);
  dft4(o, sign);
  const float s = (float)sign;
  // w1 = c*(1 + s i), w2 = s i, w3 = c*(-1 + s i)
  float2 o1 = make_float2(c * (o[1].x - s * o[1].y), c * (o[1].y + s * o[1].x));
  float2 o2 = cmul_si(o[2], sign);
  float2 o3 = make_float2(c * (-o[3].x - s * o[3].y), c * (-o[3].y + s * o[3].x));
  u[0] = cadd(e[0], o[0]);
  u[4] = csub(e[0], o[0]);
  u[1] = cadd(e[1], o1);
  u[5] = csub(e[1], o1);
  u[2] = cadd(e[2], o2);
  u[6] = csub(e[2], o2);
  u[3] = cadd(e[3], o3);
  u[7] = csub(e[3], o3);
}

}  // namespace clarity
