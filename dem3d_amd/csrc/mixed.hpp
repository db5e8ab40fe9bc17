// mixed.hpp -- the FP32 side of the mixed-precision mode (DEM_MIXED).
//
// In the mixed mode a contact is evaluated in FP64 exactly as in the FP64
// mode (CommonCalculations, src/MForces.f90:339-387); what runs in FP32 is the
// torque's cross product and the per-particle sums of the contributions,
// which are compensated (Kahan) FP32 sums added to the FP64 gravity and wall
// terms at the end.  Less than that is not accurate enough: in a sliding
// contact the dashpot term Yt (VijT - VT0') subtracts two nearly equal
// tangential velocities whose size can be hundreds of times the normal
// force, and at the end of a contact the normal force Kn overlap + Yn vn is a
// near cancellation of spring and dashpot that the Coulomb cap then carries
// into the tangential force and torque -- in FP32 either gives errors far
// above 1e-5 of the contact's force.  The stored history stays FP64.
#pragma once
#pragma clang fp contract(off)

#include "physics.hpp"

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mkf(float a, float b, float c) { return f3{a, b, c}; }
__device__ __forceinline__ f3 addf(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 subf(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 sclf(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dotf(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 crossf(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 tof(d3 a) { return f3{(float)a.x, (float)a.y, (float)a.z}; }
__device__ __forceinline__ d3 tod(f3 a) { return d3{(double)a.x, (double)a.y, (double)a.z}; }

// Compensated sum of 3-vectors (Kahan): s holds the sum, c minus the bits
// lost so far; the value is s - c.
struct KahanF3 {
  f3 s, c;
};
__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}
__device__ __forceinline__ void kahan_add3(KahanF3& k, f3 v) {
  kahan_add(k.s.x, k.c.x, v.x);
  kahan_add(k.s.y, k.c.y, v.y);
  kahan_add(k.s.z, k.c.z, v.z);
}
__device__ __forceinline__ d3 kahan_value(const KahanF3& k) {
  return d3{(double)k.s.x - (double)k.c.x, (double)k.s.y - (double)k.c.y,
            (double)k.s.z - (double)k.c.z};
}
