// physics.hpp -- per-contact arithmetic of the reference, for the kernels and the host.
//
// Every expression keeps the reference's association and the file is compiled
// with floating-point contraction off, so each multiply and add rounds on its
// own as in the reference (src/MForces.f90, src/MConfig.f90).
#pragma once
#pragma clang fp contract(off)

#include "common.hpp"

// Host and device: the host-side particle-number changes (pouring tests
// against walls, cell assignment) use the same arithmetic as the kernels.
#define DEM_HD __host__ __device__ __forceinline__

struct d3 {
  double x, y, z;
};

DEM_HD d3 mk3(double a, double b, double c) { return d3{a, b, c}; }
DEM_HD d3 add3(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DEM_HD d3 sub3(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEM_HD d3 scl3(double s, d3 a) { return d3{s * a.x, s * a.y, s * a.z}; }
DEM_HD d3 neg3(d3 a) { return d3{-a.x, -a.y, -a.z}; }
DEM_HD d3 div3(d3 a, double s) { return d3{a.x / s, a.y / s, a.z / s}; }
// dot_product accumulates from zero in index order (src/MDimension3D.f90:75)
DEM_HD double dot3(d3 a, d3 b) {
  double s = 0.0;
  s = s + a.x * b.x;
  s = s + a.y * b.y;
  s = s + a.z * b.z;
  return s;
}
// CrossProduct (src/MDimension3D.f90:83-95)
DEM_HD d3 cross3(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
DEM_HD d3 vget(const double* p) { return d3{p[0], p[1], p[2]}; }

// Fortran nint of real(8) to default integer, back to real(8): half away
// from zero; "+ 0.0" turns the -0.0 that round() gives in (-0.5, 0] into the
// +0.0 of the integer round trip.
DEM_HD double fnint(double x) { return __builtin_round(x) + 0.0; }

// Minimum image of one component (src/MForces.f90:212-214).
DEM_HD double min_image(double r, double l, double invl) {
  return r - l * fnint(r * invl);
}

DEM_HD d3 min_image3(const SysConst& s, d3 r) {
  if (s.pbc[0]) r.x = min_image(r.x, s.lbox[0], s.invl[0]);
  if (s.pbc[1]) r.y = min_image(r.y, s.lbox[1], s.invl[1]);
  if (s.pbc[2]) r.z = min_image(r.z, s.lbox[2], s.invl[2]);
  return r;
}

// One coordinate of CurrentCell3D (src/MCells.f90:507-509):
// max(0, min(int(n_by_l*x + n_by2), n-1)), with the float->int conversion
// done only on in-range values.
DEM_HD int cell_coord(double x, double n_by_l, double n_by2, int n) {
  double t = n_by_l * x + n_by2;
  if (!(t >= 1.0)) return 0;  // also catches NaN
  if (t >= (double)n) return n - 1;
  return (int)t;
}

// CurrentCell3D, 0-based linear index, x fastest; with rc_pack also the
// coordinates packed 10 bits each (x | y << 10 | z << 20), so the role and
// adjacency tests of a pair need no integer division.
struct RefCell {
  int lin, xyz;
};
DEM_HD RefCell ref_cell(const SysConst& s, double x, double y, double z) {
  int ix = cell_coord(x, s.rn_by_l[0], s.rn_by2[0], s.rn[0]);
  int iy = cell_coord(y, s.rn_by_l[1], s.rn_by2[1], s.rn[1]);
  int iz = cell_coord(z, s.rn_by_l[2], s.rn_by2[2], s.rn[2]);
  return RefCell{ix + (iy + iz * s.rn[1]) * s.rn[0], s.rc_pack ? (ix | (iy << 10) | (iz << 20)) : 0};
}

// Neighbour-search cell coordinates.  Periodic coordinates are wrapped
// first so that a particle slightly outside the box (between Euler sub-steps,
// which wrap only at the end, src/MDEM.f90:279) lands in the cell its minimum
// images see.
struct Int3 {
  int x, y, z;
};
DEM_HD Int3 grid_coords(const SysConst& s, double x, double y, double z) {
  if (s.pbc[0]) x = min_image(x, s.lbox[0], s.invl[0]);
  if (s.pbc[1]) y = min_image(y, s.lbox[1], s.invl[1]);
  if (s.pbc[2]) z = min_image(z, s.lbox[2], s.invl[2]);
  return Int3{cell_coord(x, s.g_by_l[0], s.g_by2[0], s.gn[0]),
              cell_coord(y, s.g_by_l[1], s.g_by2[1], s.gn[1]),
              cell_coord(z, s.g_by_l[2], s.g_by2[2], s.gn[2])};
}

// Key of a neighbour-search cell: 8x8x8-cell bricks in x-fastest order,
// Morton order inside a brick.  Particles are kept sorted by key, so the 27
// cells around a particle are close to it in memory in all three directions
// (with x-fastest cells, a z neighbour is a whole xy layer away).
DEM_HD int morton3_3bit(int x, int y, int z) {
  auto spread = [](int v) { return (v & 1) | ((v & 2) << 2) | ((v & 4) << 4); };
  return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}
DEM_HD int grid_key(const SysConst& s, int x, int y, int z) {
  const int b = ((z >> 3) * s.gb[1] + (y >> 3)) * s.gb[0] + (x >> 3);
  return (b << 9) | morton3_3bit(x & 7, y & 7, z & 7);
}
DEM_HD int grid_cell(const SysConst& s, double x, double y, double z) {
  const Int3 g = grid_coords(s, x, y, z);
  return grid_key(s, g.x, g.y, g.z);
}

// Signed offset from cell coordinate a to b along one reference-grid axis,
// 2 if not adjacent.  Periodic axes wrap (VerifyPBC/ApplyPBC,
// src/MCells.f90:1030-1052); the library requires >= 3 cells there.
DEM_HD int axis_offset(int a, int b, int n, int periodic) {
  int d = b - a;
  if (periodic) {
    if (d == n - 1) d = -1;
    else if (d == 1 - n) d = 1;
  }
  return (d >= -1 && d <= 1) ? d : 2;
}

struct CellXYZ {
  int x, y, z;
};
DEM_HD CellXYZ ref_cell_xyz(const SysConst& s, int c, int packed) {
  if (s.rc_pack) return CellXYZ{packed & 1023, (packed >> 10) & 1023, packed >> 20};
  int nxy = s.rn[0] * s.rn[1];
  int z = c / nxy;
  int r = c - z * nxy;
  int y = r / s.rn[0];
  return CellXYZ{r - y * s.rn[0], y, z};
}

// Are two reference cells the same or neighbours (i.e. does the reference's
// cell loop visit the pair)?  c = linear index, x = packed coordinates.
DEM_HD bool ref_adjacent(const SysConst& s, int ca, int xa, int cb, int xb) {
  if (ca == cb) return true;
  CellXYZ a = ref_cell_xyz(s, ca, xa), b = ref_cell_xyz(s, cb, xb);
  return axis_offset(a.x, b.x, s.rn[0], s.pbc[0]) != 2 &&
         axis_offset(a.y, b.y, s.rn[1], s.pbc[1]) != 2 &&
         axis_offset(a.z, b.z, s.rn[2], s.pbc[2]) != 2;
}

// Which particle of a pair plays "i" in ComputeAccelerations
// (src/MForces.f90:173-202): in one cell, the one nearer the head of the
// cell's list (NeighbourParticles, src/MCells.f90:793-829; lists are
// head-pushed, so larger entry stamp first, then larger id); across cells,
// the one whose cell lists the other cell among its 13 half-stencil cells
// (kx + Nx*(ky + Ny*kz) > 0, src/MCells.f90:454-456).
// Returns 1 if p is "i", 0 if q is, -1 if the cells are not adjacent.
DEM_HD int ref_role(const SysConst& s, int cp, int xp, int cq, int xq, int sp,
                                        int sq, int idp, int idq) {
  if (cp == cq) {
    if (sp != sq) return sp > sq;
    return idp > idq;
  }
  CellXYZ a = ref_cell_xyz(s, cp, xp), b = ref_cell_xyz(s, cq, xq);
  int dx = axis_offset(a.x, b.x, s.rn[0], s.pbc[0]);
  int dy = axis_offset(a.y, b.y, s.rn[1], s.pbc[1]);
  int dz = axis_offset(a.z, b.z, s.rn[2], s.pbc[2]);
  if (dx == 2 || dy == 2 || dz == 2) return -1;
  return (dz > 0) || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0)));
}

// ChangePlane (src/MForces.f90:391-406)
DEM_HD d3 change_plane(d3 v, d3 n) {
  d3 p = sub3(v, scl3(dot3(v, n), n));
  double nsq = dot3(p, p);
  if (nsq != 0.0) return scl3(sqrt(dot3(v, v) / nsq), p);
  return d3{0.0, 0.0, 0.0};
}

struct ContactIn {
  d3 nij, vij, vijc, disp_n, ft0, vt0;
  double delta_dsq, half_dt;
  bool contacting;
};
struct ContactOut {
  d3 fn, ft, vt;
};

// CommonCalculations (src/MForces.f90:339-387), without the accumulation
// into iForce/iTorque (done by the caller).
DEM_HD ContactOut common_calc(const ContactIn& c, d3 rij, double kn, double yn,
                                                  double kt, double yt, double musq) {
  ContactOut o;
  d3 vijn = scl3(dot3(c.vijc, c.nij), c.nij);
  o.vt = sub3(c.vijc, vijn);
  o.fn = add3(scl3(kn, c.disp_n), scl3(yn, vijn));
  if (c.contacting) {
    d3 vt0 = change_plane(c.vt0, c.nij);
    d3 ddt = scl3(c.half_dt, add3(vt0, o.vt));
    d3 dvt = sub3(o.vt, vt0);
    o.ft = add3(add3(change_plane(c.ft0, c.nij), scl3(kt, ddt)), scl3(yt, dvt));
  } else {
    double vsq = dot3(c.vij, c.vij);
    double pdt;
    if (vsq > 0.0) {
      double rv = dot3(rij, c.vij);
      double t = (rv + sqrt(rv * rv + vsq * c.delta_dsq)) / vsq;
      double two_half = 2.0 * c.half_dt;
      pdt = t < two_half ? t : two_half;
    } else {
      pdt = 0.0;
    }
    o.ft = scl3(kt * pdt + yt, o.vt);
  }
  double mu_fn_sq = musq * dot3(o.fn, o.fn);
  double ft_sq = dot3(o.ft, o.ft);
  if (ft_sq > mu_fn_sq) o.ft = scl3(sqrt(mu_fn_sq / ft_sq), o.ft);
  return o;
}

// Overlap (src/MConfig.f90:692-822).  ov = 1 overlap, 0 none, -1 the
// degenerate cone case that needs the random generator (src/MConfig.f90:808).
struct WallHit {
  int ov;
  double dij;
  d3 nij, qji;
};

DEM_HD WallHit wall_overlap(const SysConst& s, const WallGeom& w, d3 xi,
                                                double ri, double max_alpha_u, double min_aw_sq,
                                                double max_aw_sq) {
  WallHit h;
  h.ov = 0;
  h.dij = 0.0;
  h.nij = mk3(0.0, 0.0, 0.0);
  h.qji = h.nij;
  const d3 rij = min_image3(s, sub3(vget(w.pos), xi));
  const d3 U = vget(w.u), W = vget(w.w), N = vget(w.n);
  if (w.shape == 1) {
    const double an = dot3(rij, N);
    const double abs_n = fabs(an);
    if (!(abs_n < ri)) return h;
    double au = dot3(rij, U), aw = dot3(rij, W);
    const double abs_u = fabs(au), abs_w = fabs(aw);
    if (!((abs_u < w.l1 + ri) && (abs_w < w.l2 + ri))) return h;
    if ((abs_u < w.l1) && (abs_w < w.l2)) {
      h.nij = scl3(copysign(1.0, an), N);
      h.dij = abs_n;
      h.qji = neg3(add3(scl3(au, U), scl3(aw, W)));
      h.ov = 1;
      return h;
    }
    au = fmax(-w.l1, fmin(au, w.l1));
    aw = fmax(-w.l2, fmin(aw, w.l2));
    h.qji = neg3(add3(scl3(au, U), scl3(aw, W)));
    const d3 nn = add3(rij, h.qji);
    const double dsq = dot3(nn, nn);
    if (!(dsq < ri * ri)) return h;
    h.dij = sqrt(dsq);
    h.nij = div3(nn, h.dij);
    h.ov = 1;
    return h;
  }
  // cone
  const double au = dot3(rij, U);
  if (!(fabs(au) < max_alpha_u)) return h;
  const d3 rw = sub3(rij, scl3(au, U));
  const double awsq = dot3(rw, rw);
  if (!((awsq >= min_aw_sq) && (awsq < max_aw_sq))) return h;
  const double aw = sqrt(awsq);
  const double lby2_au = w.l1 - au;
  const double ra_aw = w.l2 - aw;
  double lambda = lby2_au * w.sin_t - ra_aw * w.cos_t;
  if ((lambda >= 0.0) && (lambda <= w.rap)) {
    const double sd = -(lby2_au * w.cos_t + ra_aw * w.sin_t);
    h.dij = fabs(sd);
    if (!(h.dij < ri)) return h;
    if (aw == 0.0) { h.ov = -1; return h; }
    const d3 wv = div3(rw, aw);
    const d3 nn = add3(scl3(w.sin_t, wv), scl3(w.cos_t, U));
    h.qji = sub3(scl3(sd, nn), rij);
    h.nij = div3(add3(h.qji, rij), h.dij);
    h.ov = 1;
    return h;
  }
  lambda = fmax(0.0, fmin(lambda, w.rap));
  if (aw == 0.0) { h.ov = -1; return h; }
  const d3 wv = div3(rw, aw);
  const d3 nn = sub3(scl3(w.sin_t * lambda - lby2_au, U), scl3(w.cos_t * lambda + ra_aw, wv));
  const double dsq = dot3(nn, nn);
  if (!(dsq < ri * ri)) return h;
  h.dij = sqrt(dsq);
  h.qji = sub3(nn, rij);
  h.nij = div3(nn, h.dij);
  h.ov = 1;
  return h;
}
