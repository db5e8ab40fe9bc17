// mixed.hpp -- the pair contact of the mixed-precision mode (DEM_MIXED).
//
// What stays FP64: the pair separation and the overlap test
// (MinDijSq - DijSq > 0, src/MForces.f90:223-224), so contact sets are the
// FP64 ones; the normal overlap Dij - MinDij (a small difference of large
// numbers); and the stored tangential spring FT -- its rotation into the new
// tangent plane (ChangePlane, src/MForces.f90:391-406), its accumulation
// (src/MForces.f90:356-358) and its Coulomb rescale (src/MForces.f90:376-379)
// -- because it integrates over the whole life of a contact; and the time
// since the first touch of a new contact (a cancelling root).  Everything
// else of CommonCalculations (src/MForces.f90:339-387) -- normal, relative
// velocities, normal force, the increment of the spring, the stored tangential
// velocity VT, torques -- is FP32, and the per-particle sums of the FP32 pair
// forces are compensated (Kahan) FP32 sums.
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

// ChangePlane in FP32 (for the stored tangential velocity).
__device__ __forceinline__ f3 change_plane_f(f3 v, f3 n) {
  const f3 p = subf(v, sclf(dotf(v, n), n));
  const float nsq = dotf(p, p);
  if (nsq != 0.0f) return sclf(sqrtf(dotf(v, v) / nsq), p);
  return f3{0.0f, 0.0f, 0.0f};
}

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

struct MixedOut {
  f3 fn, ft, tq_dir;  // normal force, tangential force (as applied), Nij x FT
  f3 n;               // unit normal i -> j
  d3 ft_store;        // FP64 spring to store
  f3 vt;              // tangential velocity to store
};

// One pair contact, seen from particle i ("i" in the frame of the caller;
// ft0/vt0 already carry the orientation sign).  rij = xj - xi (minimum
// image, FP64), dsq = |rij|^2, delta_dsq = MinDijSq - dsq > 0, dv = vj - vi
// (FP64 difference).
__device__ __forceinline__ MixedOut contact_mixed(bool contacting, d3 rij, double dsq,
                                                  double delta_dsq, double mindij, d3 dv, d3 wi,
                                                  d3 wj, double ri, double rj, d3 ft0, d3 vt0,
                                                  float half_dt, float kn, float yn, float kt,
                                                  float yt, double musq) {
  MixedOut o;
  const double dij = sqrt(dsq);
  const float inv = 1.0f / (float)dij;
  const f3 n = sclf(inv, tof(rij));
  o.n = n;
  const f3 disp_n = sclf((float)(dij - mindij), n);
  const f3 vij = tof(dv);
  const f3 wsum = addf(sclf((float)ri, tof(wi)), sclf((float)rj, tof(wj)));
  const f3 vijc = subf(vij, crossf(wsum, n));
  const f3 vijn = sclf(dotf(vijc, n), n);
  o.vt = subf(vijc, vijn);
  o.fn = addf(sclf(kn, disp_n), sclf(yn, vijn));
  d3 ft;
  if (contacting) {
    const f3 vt0r = change_plane_f(tof(vt0), n);
    const d3 ft0r = change_plane(ft0, tod(n));
    const f3 ddt = sclf(half_dt, addf(vt0r, o.vt));
    const f3 dvt = subf(o.vt, vt0r);
    ft = add3(ft0r, tod(addf(sclf(kt, ddt), sclf(yt, dvt))));
  } else {
    // Time since the first touch (src/MForces.f90:360-370) in FP64: for an
    // approaching pair rv < 0 and vsq*delta << rv^2, so rv + sqrt(rv^2 +
    // vsq*delta) cancels and FP32 would keep few digits of it.  Only new
    // contacts take this branch.
    const double vsq = dot3(dv, dv);
    double pdt = 0.0;
    if (vsq > 0.0) {
      const double rv = dot3(rij, dv);
      const double t = (rv + sqrt(rv * rv + vsq * delta_dsq)) / vsq;
      const double two_half = 2.0 * (double)half_dt;
      pdt = t < two_half ? t : two_half;
    }
    ft = tod(sclf(kt * (float)pdt + yt, o.vt));
  }
  const double mu_fn_sq = musq * (double)dotf(o.fn, o.fn);
  const double ft_sq = dot3(ft, ft);
  if (ft_sq > mu_fn_sq) ft = scl3(sqrt(mu_fn_sq / ft_sq), ft);
  o.ft_store = ft;
  o.ft = tof(ft);
  o.tq_dir = crossf(n, o.ft);
  return o;
}
