// files.hpp -- configuration (.cfg) and state (.stt) files in the
// reference's binary layout, host code only.
//
// BinFiles (src/BinFiles.f90) writes every value as its raw bytes through a
// 1 KiB buffer flushed in direct-access records, so a file is the plain
// concatenation of the values written: integers and reals little-endian at
// their declared kinds, character variables without a length, logicals at
// their kinds (Finished is logical(2), the Vectors flag logical(1)).
//
//   configuration  SaveConfToUnit / ReadConfFromUnit (src/MConfig.f90:109-448)
//   state file     SaveState / ReadState (src/DEM.f90:947-1110): run records,
//                  configuration (Complete), SaveCellStructure
//                  (src/MCells.f90:960-983), SaveIntegrationVariables
//                  (src/MDEM.f90:52-68), SaveContactList
//                  (src/MContacts.f90:75-118), random-number state, Conf,
//                  Step + 1, Ncol.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dem3d.h"

namespace demfile {

struct Writer {
  std::vector<unsigned char> b;
  void raw(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    b.insert(b.end(), c, c + n);
  }
  void i8(int v) { const int8_t x = (int8_t)v; raw(&x, 1); }
  void i16(int v) { const int16_t x = (int16_t)v; raw(&x, 2); }
  void i32(int32_t v) { raw(&v, 4); }
  void f64(double v) { raw(&v, 8); }
  void f64s(const double* v, int n) { raw(v, 8 * (size_t)n); }
  void chars(const char* s, size_t n) {
    // character(n): blank-padded when the given string is shorter
    size_t k = 0;
    while (k < n && s[k]) ++k;
    raw(s, k);
    for (; k < n; ++k) b.push_back((unsigned char)' ');
  }
};

struct Reader {
  std::vector<unsigned char> b;
  size_t off = 0;
  bool ok = true;
  bool take(void* p, size_t n) {
    if (!ok || off + n > b.size()) {
      ok = false;
      if (p) memset(p, 0, n);
      return false;
    }
    if (p) memcpy(p, b.data() + off, n);
    off += n;
    return true;
  }
  int i8() { int8_t x = 0; take(&x, 1); return x; }
  int i16() { int16_t x = 0; take(&x, 2); return x; }
  int32_t i32() { int32_t x = 0; take(&x, 4); return x; }
  double f64() { double x = 0; take(&x, 8); return x; }
  void f64s(double* v, int n) { take(v, 8 * (size_t)n); }
  void skip(size_t n) { take(nullptr, n); }
};

inline bool save(const char* path, const Writer& w) {
  FILE* f = fopen(path, "wb");
  if (!f) return false;
  const bool good = fwrite(w.b.data(), 1, w.b.size(), f) == w.b.size();
  return fclose(f) == 0 && good;
}

inline bool load(const char* path, Reader& r) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  unsigned char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) r.b.insert(r.b.end(), buf, buf + n);
  const bool err = ferror(f) != 0;
  fclose(f);
  return !err;
}

// CartToSpheric and SphericToCart (src/MDimension3D.f90:151-207).
inline void cart_to_spheric(const double* u, double* phi) {
  const double pi_by2 = 1.57079632679489662;
  phi[1] = std::asin(u[2]);
  const double cos_phi2 = std::cos(phi[1]);
  if (cos_phi2 == 0.0) {
    phi[0] = 0.0;
  } else {
    const double aux = 1.0 / cos_phi2;
    if (aux * u[0] >= 0.0) phi[0] = std::asin(aux * u[1]);
    else phi[0] = std::acos(aux * u[1]) + pi_by2;
  }
}

inline void spheric_to_cart(const double* phi, double* u) {
  const double cos_phi2 = std::cos(phi[1]);
  u[0] = std::cos(phi[0]) * cos_phi2;
  u[1] = std::sin(phi[0]) * cos_phi2;
  u[2] = std::sin(phi[1]);
}

// GramSchmidt (src/MUtil.f90:163-171) and U .x. W (src/MDimension3D.f90:86-95).
inline void gram_schmidt(double* u, double* w) {
  const double nu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int d = 0; d < 3; ++d) u[d] = u[d] / nu;
  const double uw = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
  for (int d = 0; d < 3; ++d) w[d] = w[d] - uw * u[d];
  const double nw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (int d = 0; d < 3; ++d) w[d] = w[d] / nw;
}

inline void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// int(x, 2) of Fortran: truncation toward zero to a 16-bit integer.
inline int to_i16(double x) { return (int)(int16_t)(int32_t)std::trunc(x); }

inline double max_lbox(const double* l) {
  double m = l[0];
  if (l[1] > m) m = l[1];
  if (l[2] > m) m = l[2];
  return m;
}

// SaveConfToUnit (src/MConfig.f90:109-236).  Returns an empty string or
// the reason the configuration cannot be written.
inline std::string write_config(Writer& o, const dem_config_t& c) {
  const bool full = c.form == DEM_CFG_COMPLETE;
  if (!full && c.form != DEM_CFG_COMPACT) return "form must be DEM_CFG_COMPLETE or DEM_CFG_COMPACT";
  if (c.nc < 0 || c.nw < 0 || c.np < 0) return "negative count";
  if ((c.nc && (!c.number || !c.radius)) || (c.nw && !c.walls) ||
      (c.np && (!c.comp || !c.x)))
    return "missing array";
  if (full && c.nc && (!c.density || !c.stiffness || !c.shear_stiffness || !c.mass || !c.inertia))
    return "missing component array";
  if (full && c.np && (!c.v || !c.w)) return "missing velocity array";
  std::vector<int32_t> count((size_t)c.nc, 0);
  for (int i = 0; i < c.np; ++i) {
    if (c.comp[i] < 1 || c.comp[i] > c.nc) return "particle with an invalid component";
    count[c.comp[i] - 1] += 1;
  }
  for (int k = 0; k < c.nc; ++k)
    if (count[k] != c.number[k]) return "number[] does not count the particles of each component";
  const double highest = 32766.0;  // real(huge(1_2)-1)
  const double factor = 2.0 * highest / max_lbox(c.lbox);
  const double a = 0.8 * highest / std::acos(-1.0);
  const double bb = -0.6 * highest;
  const size_t N = (size_t)c.np;
  o.raw("DEMC", 4);
  o.i8(2);  // Version
  o.i8(3);  // dim
  o.i8(c.form);
  o.f64(c.length_scale);
  o.f64(c.time_scale);
  if (full) o.f64(c.mass_scale);
  o.f64(c.time);
  o.f64s(c.lbox, 3);
  o.i32(c.nc);
  for (int k = 0; k < c.nc; ++k) {
    o.i32(c.number[k]);
    o.f64(c.radius[k]);
    if (full) {
      o.f64(c.density[k]);
      o.f64(c.stiffness[k]);
      o.f64(c.shear_stiffness[k]);
      o.f64(c.mass[k]);
      o.f64(c.inertia[k]);
    }
  }
  o.i32(c.nw);
  for (int j = 0; j < c.nw; ++j) {
    const dem_wall_t& w = c.walls[j];
    o.i32(w.shape);
    o.f64(w.l1);
    o.f64(w.l2);
    o.f64(w.l3);
    if (full) {
      o.f64s(w.u, 3);
      o.f64s(w.w, 3);
      o.f64s(w.n, 3);
    } else {
      double phi[2];
      cart_to_spheric(w.u, phi);
      o.i16(to_i16(a * phi[0] + bb));
      o.i16(to_i16(a * phi[1] + bb));
      cart_to_spheric(w.w, phi);
      o.i16(to_i16(a * phi[0] + bb));
      o.i16(to_i16(a * phi[1] + bb));
    }
    o.f64s(w.position, 3);
    if (full) {
      o.f64s(w.velocity, 3);
      o.f64s(w.omega, 3);
      o.f64(w.stiffness);
      o.f64(w.shear_stiffness);
    }
  }
  o.i8(0);  // Vectors = allocated(Direction): no orientation vectors
  if (full) {
    for (size_t i = 0; i < N; ++i) {
      o.i32(c.comp[i]);
      for (int d = 0; d < 3; ++d) o.f64(c.x[d * N + i]);
      for (int d = 0; d < 3; ++d) o.f64(c.v[d * N + i]);
      for (int d = 0; d < 3; ++d) o.f64(c.w[d * N + i]);
    }
  } else {
    for (int k = 1; k <= c.nc; ++k)
      for (size_t i = 0; i < N; ++i)
        if (c.comp[i] == k)
          for (int d = 0; d < 3; ++d) o.i16(to_i16(factor * c.x[d * N + i]));
  }
  return "";
}

// ReadConfFromUnit (src/MConfig.f90:266-448).  Sets the counts and scalars
// of c, fills its non-NULL arrays.
inline std::string read_config(Reader& r, dem_config_t& c) {
  char sig[4];
  r.take(sig, 4);
  if (!r.ok || memcmp(sig, "DEMC", 4) != 0) return "not a configuration file";
  const int version = r.i8();
  const int dim = r.i8();
  if (dim != 3) return "not a three-dimensional configuration";
  const int form = r.i8();
  const bool full = form == DEM_CFG_COMPLETE;
  c.form = form;
  c.version = version;
  c.length_scale = r.f64();
  c.time_scale = r.f64();
  c.mass_scale = full ? r.f64() : 0.0;
  c.time = r.f64();
  r.f64s(c.lbox, 3);
  const double highest = 32766.0;
  const double pi = std::acos(-1.0);
  const double a = 1.25 * pi / highest;
  const double bb = 0.75 * pi;
  const double factor = 0.5 * max_lbox(c.lbox) / highest;
  c.nc = r.i32();
  if (!r.ok || c.nc < 0 || c.nc > DEM_MAX_COMPONENTS) return "bad component count";
  std::vector<int32_t> number((size_t)c.nc);
  for (int k = 0; k < c.nc; ++k) {
    number[k] = r.i32();
    const double rad = r.f64();
    double den = 0, st = 0, sh = 0, ms = 0, in = 0;
    if (full) {
      den = r.f64();
      st = r.f64();
      sh = version == 2 ? r.f64() : 2.0 / 7.0 * st;
      ms = r.f64();
      in = r.f64();
    }
    if (c.number) c.number[k] = number[k];
    if (c.radius) c.radius[k] = rad;
    if (c.density) c.density[k] = den;
    if (c.stiffness) c.stiffness[k] = st;
    if (c.shear_stiffness) c.shear_stiffness[k] = sh;
    if (c.mass) c.mass[k] = ms;
    if (c.inertia) c.inertia[k] = in;
  }
  c.nw = r.i32();
  if (!r.ok || c.nw < 0 || c.nw > DEM_MAX_WALLS) return "bad wall count";
  for (int j = 0; j < c.nw; ++j) {
    dem_wall_t w;
    memset(&w, 0, sizeof w);
    w.shape = r.i32();
    w.l1 = r.f64();
    w.l2 = r.f64();
    w.l3 = r.f64();
    if (full) {
      r.f64s(w.u, 3);
      r.f64s(w.w, 3);
      r.f64s(w.n, 3);
    } else {
      double phi[2];
      phi[0] = a * (double)r.i16() + bb;
      phi[1] = a * (double)r.i16() + bb;
      spheric_to_cart(phi, w.u);
      phi[0] = a * (double)r.i16() + bb;
      phi[1] = a * (double)r.i16() + bb;
      spheric_to_cart(phi, w.w);
      gram_schmidt(w.u, w.w);
      cross(w.u, w.w, w.n);
    }
    r.f64s(w.position, 3);
    if (full) {
      r.f64s(w.velocity, 3);
      r.f64s(w.omega, 3);
      if (version == 2) {
        w.stiffness = r.f64();
        w.shear_stiffness = r.f64();
      }
    }
    if (c.walls) c.walls[j] = w;
  }
  long long np = 0;
  for (int k = 0; k < c.nc; ++k) np += number[k];
  if (np < 0 || np > 0x7fffffff) return "bad particle count";
  c.np = (int32_t)np;
  const size_t N = (size_t)np;
  const bool vectors = r.i8() != 0;
  if (full) {
    for (size_t i = 0; i < N; ++i) {
      const int32_t idx = r.i32();
      double p[9];
      r.f64s(p, 9);
      if (vectors) r.skip(9 * 8);
      if (c.comp) c.comp[i] = idx;
      for (int d = 0; d < 3; ++d) {
        if (c.x) c.x[d * N + i] = p[d];
        if (c.v) c.v[d * N + i] = p[3 + d];
        if (c.w) c.w[d * N + i] = p[6 + d];
      }
    }
  } else {
    size_t i = 0;
    for (int k = 1; k <= c.nc; ++k)
      for (int q = 0; q < number[k - 1]; ++q, ++i) {
        int rr[3];
        for (int d = 0; d < 3; ++d) rr[d] = r.i16();
        if (vectors) r.skip(4 * 2);
        if (c.comp) c.comp[i] = k;
        for (int d = 0; d < 3; ++d) {
          if (c.x) c.x[d * N + i] = factor * (double)rr[d];
          if (c.v) c.v[d * N + i] = 0.0;
          if (c.w) c.w[d * N + i] = 0.0;
        }
      }
  }
  if (!r.ok) return "file ends inside the configuration";
  return "";
}

// The run records that precede the configuration in a state file
// (src/DEM.f90:956-982).
inline std::string write_run_head(Writer& o, const dem_run_t& run) {
  if (run.n_interruptions < 0) return "negative n_interruptions";
  if (run.n_interruptions > 0 && (!run.interruption || !run.retaking))
    return "missing interruption records";
  o.i16(run.finished ? 1 : 0);
  o.chars(run.start_time, 28);
  o.i32(run.n_interruptions);
  for (int k = 0; k < run.n_interruptions; ++k) {
    o.chars(run.interruption + 28 * (size_t)k, 28);
    o.chars(run.retaking + 28 * (size_t)k, 28);
  }
  o.chars(run.finish_time, 28);
  o.i32(run.hours);
  o.i8(run.minutes);
  o.i8(run.seconds);
  o.i16(run.milisecs);
  return "";
}

inline void read_run_head(Reader& r, dem_run_t& run) {
  char* inter = run.interruption;
  char* ret = run.retaking;
  run.finished = r.i16() != 0;
  r.take(run.start_time, 28);
  run.n_interruptions = r.i32();
  for (int k = 0; k < run.n_interruptions && r.ok; ++k) {
    r.take(inter ? inter + 28 * (size_t)k : nullptr, 28);
    r.take(ret ? ret + 28 * (size_t)k : nullptr, 28);
  }
  r.take(run.finish_time, 28);
  run.hours = r.i32();
  run.minutes = r.i8();
  run.seconds = r.i8();
  run.milisecs = r.i16();
}

// SaveCellStructure, SaveIntegrationVariables, SaveContactList.
inline std::string write_sections(Writer& o, const dem_state_sections_t& s, int np) {
  const size_t N = (size_t)np;
  if (s.ncell < 0 || (s.ncell && !s.cell_len) || (np && !s.cell_members))
    return "missing cell lists";
  long long listed = 0;
  for (int c = 0; c < s.ncell; ++c) listed += s.cell_len[c];
  if (listed != np) return "the cell lists do not hold every particle once";
  if (s.nprev < 0 || s.nprev > 3) return "nprev must be 0..3";
  if (s.nprev && np && !s.ring) return "missing ring";
  if ((s.npairs && (!s.keys || !s.ft || !s.vt)) || (s.nwall && (!s.wkeys || !s.wft || !s.wvt)))
    return "missing contact arrays";
  o.i32(s.ncell);
  size_t m = 0;
  for (int c = 0; c < s.ncell; ++c) {
    o.i32(s.cell_len[c]);
    for (int k = 0; k < s.cell_len[c]; ++k) o.i32(s.cell_members[m++]);
  }
  for (int j = 0; j < s.nprev; ++j) o.i32(s.ipc[j]);
  for (int j = 0; j < s.nprev; ++j)
    for (size_t i = 0; i < N; ++i)
      for (int q = 0; q < 9; ++q) o.f64(s.ring[((size_t)j * 9 + q) * N + i]);
  o.i32(np);
  int64_t kp = 0, kw = 0;
  for (int i = 1; i <= np; ++i) {
    for (int pass = 0; pass < 2; ++pass) {
      const int64_t n = pass == 0 ? s.npairs : s.nwall;
      const uint64_t* keys = pass == 0 ? s.keys : s.wkeys;
      const double* v1 = pass == 0 ? s.ft : s.wft;
      const double* v2 = pass == 0 ? s.vt : s.wvt;
      int64_t& k = pass == 0 ? kp : kw;
      int64_t e = k;
      while (e < n && (int64_t)(keys[e] >> 32) == i) ++e;
      if (e < n && (int64_t)(keys[e] >> 32) < i) return "contact keys are not ascending";
      o.i32((int32_t)(e - k));
      for (; k < e; ++k) {
        o.i32((int32_t)(keys[k] & 0xffffffffu));
        o.f64s(v1 + 3 * k, 3);
        o.f64s(v2 + 3 * k, 3);
      }
    }
  }
  if (kp != s.npairs || kw != s.nwall) return "contact keys outside 1..np or not ascending";
  return "";
}

inline std::string read_sections(Reader& r, dem_state_sections_t& s, int np) {
  const size_t N = (size_t)np;
  s.ncell = r.i32();
  if (!r.ok || s.ncell < 0) return "bad cell count";
  size_t m = 0;
  for (int c = 0; c < s.ncell && r.ok; ++c) {
    const int32_t n = r.i32();
    if (n < 0 || m + (size_t)n > N) return "cell lists hold more particles than the configuration";
    if (s.cell_len) s.cell_len[c] = n;
    for (int k = 0; k < n; ++k, ++m) {
      const int32_t q = r.i32();
      if (s.cell_members) s.cell_members[m] = q;
    }
  }
  if (r.ok && m != N) return "cell lists do not hold every particle";
  for (int j = 0; j < s.nprev; ++j) s.ipc[j] = r.i32();
  for (int j = 0; j < s.nprev; ++j)
    for (size_t i = 0; i < N; ++i) {
      double p[9];
      r.f64s(p, 9);
      if (s.ring)
        for (int q = 0; q < 9; ++q) s.ring[((size_t)j * 9 + q) * N + i] = p[q];
    }
  const int32_t npl = r.i32();
  if (r.ok && npl != np) return "contact lists are not sized for the particles";
  int64_t kp = 0, kw = 0;
  for (int i = 1; i <= npl && r.ok; ++i) {
    for (int pass = 0; pass < 2 && r.ok; ++pass) {
      const int32_t n = r.i32();
      if (n < 0) return "negative list length";
      for (int k = 0; k < n && r.ok; ++k) {
        const int32_t idx = r.i32();
        double v[6];
        r.f64s(v, 6);
        const uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)idx;
        uint64_t* keys = pass == 0 ? s.keys : s.wkeys;
        double* v1 = pass == 0 ? s.ft : s.wft;
        double* v2 = pass == 0 ? s.vt : s.wvt;
        int64_t& q = pass == 0 ? kp : kw;
        if (keys) keys[q] = key;
        if (v1) memcpy(v1 + 3 * q, v, 24);
        if (v2) memcpy(v2 + 3 * q, v + 3, 24);
        ++q;
      }
    }
  }
  s.npairs = kp;
  s.nwall = kw;
  if (!r.ok) return "file ends inside the state records";
  return "";
}

// Records after the contact list (src/DEM.f90:995-1007).
inline void write_run_tail(Writer& o, const dem_run_t& run) {
  o.f64(run.aam);
  o.i32(run.dummy);
  o.i32(run.iix);
  o.i32(run.iiy);
  o.i32(run.iik);
  o.i32(run.conf);
  o.i32(run.next_step);
  o.i32(run.ncol);
}

inline void read_run_tail(Reader& r, dem_run_t& run) {
  run.aam = r.f64();
  run.dummy = r.i32();
  run.iix = r.i32();
  run.iiy = r.i32();
  run.iik = r.i32();
  run.conf = r.i32();
  run.next_step = r.i32();
  run.ncol = r.i32();
}

}  // namespace demfile
