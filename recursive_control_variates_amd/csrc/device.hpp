// Device math for the render kernels: 3-vectors, the PCG32 generator with
// TEA seeding, sampling warps and shading frames.
//
// Every routine is written from its published description:
//   PCG32      M. E. O'Neill, "PCG: A family of simple fast space-efficient
//              statistically good algorithms for random number generation",
//              2014 (pcg32: 64-bit LCG state, XSH-RR output).
//   TEA        Wheeler & Needham, "TEA, a tiny encryption algorithm", 1994,
//              used as a per-lane seed scrambler as in Zafar, Olano & Curtis,
//              "GPU random numbers via the tiny encryption algorithm", HPG 2010.
//   concentric Shirley & Chiu, "A low distortion map between disk and square",
//              1997.
//   frame      Duff et al., "Building an orthonormal basis, revisited", JCGT 2017.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RCV_HD __host__ __device__ __forceinline__

namespace rcv {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv4Pi = 0.07957747154594766788f;
constexpr float kInf = __builtin_huge_valf();
// Ray offset scale (relative to the magnitude of the hit point) and the
// fraction of a shadow segment left out at its far end.
constexpr float kRayEps = 8.9407e-5f;
constexpr float kShadowEps = 8.9407e-4f;

struct V3 {
  float x, y, z;
};

RCV_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
RCV_HD V3 v3(float a) { return V3{a, a, a}; }
RCV_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RCV_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RCV_HD V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
RCV_HD V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
RCV_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RCV_HD V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
RCV_HD V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
RCV_HD V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
RCV_HD V3 &operator+=(V3 &a, V3 b) {
  a = a + b;
  return a;
}
RCV_HD V3 &operator-=(V3 &a, V3 b) {
  a = a - b;
  return a;
}
RCV_HD V3 &operator*=(V3 &a, V3 b) {
  a = a * b;
  return a;
}
RCV_HD V3 &operator*=(V3 &a, float s) {
  a = a * s;
  return a;
}
RCV_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RCV_HD V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RCV_HD float norm(V3 a) { return sqrtf(dot(a, a)); }
RCV_HD V3 normalize(V3 a) { return a * (1.0f / norm(a)); }
RCV_HD float hmax(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
RCV_HD V3 vmax(V3 a, V3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
RCV_HD bool any_nonzero(V3 a) { return a.x != 0.0f || a.y != 0.0f || a.z != 0.0f; }
RCV_HD float sqr(float a) { return a * a; }

// ------------------------------------------------------------------ RNG
RCV_HD void tea(uint32_t &v0, uint32_t &v1, int rounds = 4) {
  uint32_t sum = 0;
  for (int i = 0; i < rounds; ++i) {
    sum += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
  }
}

RCV_HD uint64_t tea64(uint32_t a, uint32_t b) {
  tea(a, b);
  return (uint64_t(b) << 32) | uint64_t(a);
}

struct Pcg32 {
  uint64_t state;
  uint64_t inc;

  RCV_HD uint32_t next_u32() {
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
  }
  RCV_HD void seed(uint64_t initstate, uint64_t initseq) {
    state = 0;
    inc = (initseq << 1u) | 1u;
    next_u32();
    state += initstate;
    next_u32();
  }
  // Uniform float in [0, 1): 23 random mantissa bits.
  RCV_HD float next_1d() {
    const uint32_t bits = (next_u32() >> 9) | 0x3f800000u;
    return __builtin_bit_cast(float, bits) - 1.0f;
  }
};

// Per camera-sample generator: depends only on (seed, lane) where lane =
// pixel * spp + sample, never on the launch shape.
RCV_HD Pcg32 lane_rng(uint32_t seed, uint32_t lane) {
  Pcg32 r;
  r.seed(tea64(seed, lane), tea64(lane, seed));
  return r;
}

// ------------------------------------------------------------------ warps
RCV_HD V3 square_to_uniform_sphere(float u, float v) {
  const float z = 1.0f - 2.0f * u;
  const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
  float s, c;
  sincosf(2.0f * kPi * v, &s, &c);
  return {r * c, r * s, z};
}

RCV_HD void square_to_disk_concentric(float u, float v, float &x, float &y) {
  const float a = 2.0f * u - 1.0f, b = 2.0f * v - 1.0f;
  if (a == 0.0f && b == 0.0f) {
    x = y = 0.0f;
    return;
  }
  float r, phi;
  if (fabsf(a) > fabsf(b)) {
    r = a;
    phi = (kPi * 0.25f) * (b / a);
  } else {
    r = b;
    phi = kPi * 0.5f - (kPi * 0.25f) * (a / b);
  }
  float s, c;
  sincosf(phi, &s, &c);
  x = r * c;
  y = r * s;
}

RCV_HD V3 square_to_cosine_hemisphere(float u, float v) {
  float x, y;
  square_to_disk_concentric(u, v, x, y);
  const float z = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
  return {x, y, z};
}

// Orthonormal frame around a unit normal n.
struct Frame {
  V3 s, t, n;
  RCV_HD V3 to_local(V3 v) const { return {dot(v, s), dot(v, t), dot(v, n)}; }
  RCV_HD V3 to_world(V3 v) const { return s * v.x + t * v.y + n * v.z; }
};

RCV_HD Frame make_frame(V3 n) {
  const float sign = copysignf(1.0f, n.z);
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  Frame f;
  f.s = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
  f.t = {b, sign + n.y * n.y * a, -n.y};
  f.n = n;
  return f;
}

// Power heuristic with exponent 2 (plugins/twostatepath.py:6 uses the
// upstream mis_weight).
RCV_HD float mis_weight(float pdf_a, float pdf_b) {
  const float a2 = pdf_a * pdf_a;
  return pdf_a > 0.0f ? a2 / (pdf_b * pdf_b + a2) : 0.0f;
}

// Offsets a surface point along its normal towards direction d.
RCV_HD V3 offset_point(V3 p, V3 n, V3 d) {
  float mag = (1.0f + hmax(v3(fabsf(p.x), fabsf(p.y), fabsf(p.z)))) * kRayEps;
  mag = copysignf(mag, dot(n, d));
  return p + n * mag;
}

}  // namespace rcv
