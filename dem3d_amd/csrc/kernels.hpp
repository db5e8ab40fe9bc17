// kernels.hpp -- gfx950 kernels of the DEM per-timestep path.
//
// Layout in HBM (internal particle order = cell order of the last resort):
//   pos   double4[N]   x, y, z, component (0-based, as a double)
//   vw    double[6N]   48-byte records vx vy vz wx wy wz
//   ring  double[nprev][9][N]  ABM ring Previous(:,i) (src/MDEM.f90:14-24)
//   nbr   int[max_nb][N]       directed neighbour list, slot-major (ELL)
//   contact table (two copies, swapped by every Update evaluation):
//         cnt int[N], part int[max_c][N], hist double[max_c][6][N]
//   wall history: wflag uint[N] (bit w = contact with wall w), whist [nw][6][N]
// A wavefront owns 64 consecutive particles; every particle accumulates its
// own force and torque in registers (the reference's scattered Force(j,:)
// updates, src/MForces.f90:256-259, become a per-particle sum over its
// directed list: atomic-free and run-to-run deterministic).
#pragma once
#pragma clang fp contract(off)

#include "physics.hpp"

enum { MODE_ACC = 0, MODE_EULER = 1, MODE_PRED = 2, MODE_CORR = 3 };

// Blocks are dispatched round-robin over the 8 XCDs; map consecutive block
// ranges to one XCD so a spatial slab of particles shares one L2.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7, l = b >> 3;
  return (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_or(int v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

struct PassArgs {
  // evaluated positions / velocities (Particle or Predicted)
  const double4* pos_in;
  const double* vw_in;
  // Particle state (read by PRED/CORR/EULER; written by CORR; EULER writes *_out)
  double4* pos;
  double* vw;
  double4* pos_out;     // PRED: predicted; EULER: next Particle buffer
  double* vw_out;
  double* acc_out;      // ACC/EULER: a, alpha of the evaluated set; CORR: predicted accel (optional)
  double* ring;
  const double* slot_rd[3];  // ring slots in ipc order (k = 1..nprev)
  double* slot_wr;           // PRED: slot that becomes ipc(1)
  double coef0, coef[3];     // AB0/AB(k) (PRED), AM0/AM(k) (CORR)
  double ab0, ab[3];         // CORR: AB coefficients for the look-ahead predictor
  double h, hsq2;            // EULER sub-step and 0.5*h**2
  double half_dt;            // HalfTimeStep
  const int* id;
  const int* rc_in;
  const int* st_in;
  int* rc_out;
  int* st_out;
  int stamp_next;
  int check_pred;            // PRED: verify the predicted positions against xref
  int step_idx;              // CORR: step index within the current dem_step call
  int par;                   // evaluation parity (counters)
  const double4* xref;
  const int* nbr;
  const int* nnb;
  const int* ccnt_in;
  const int* cpart_in;
  const double* chist_in;
  int* ccnt_out;
  int* cpart_out;
  double* chist_out;
  unsigned* wflag;
  double* whist;
  int* rb_clear;             // rebuild flag consumed before this evaluation
  int* rb_set;               // rebuild flag for the next evaluation
  Ctl* ctl;
};

template <int MODE, bool UPDATE>
__global__ __launch_bounds__(DEM_BLOCK) void k_pass(SysConst s, PassArgs a) {
  Ctl* ctl = a.ctl;
  if (*(volatile int*)&ctl->escaped) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *a.rb_clear = 0;
    ctl->pairs[a.par ^ 1] = 0;
    ctl->wallc[a.par ^ 1] = 0;
    if (UPDATE) ctl->stored[a.par ^ 1] = 0;
  }
  const size_t N = (size_t)s.np;
  const int i = xcd_block() * DEM_BLOCK + threadIdx.x;
  unsigned long long n_col = 0, n_pair = 0, n_wall = 0, n_stored = 0;
  int err = 0, need_rb = 0, n_esc = 0, cnt_hw = 0;
  if (i < s.np) {
    const double4 P = a.pos_in[i];
    const int ci = (int)P.w;
    const d3 xi = mk3(P.x, P.y, P.z);
    const double2* vwp = reinterpret_cast<const double2*>(a.vw_in + 6 * (size_t)i);
    const double2 u0 = vwp[0], u1 = vwp[1], u2 = vwp[2];
    const d3 vi = mk3(u0.x, u0.y, u1.x);
    const d3 wi = mk3(u1.y, u2.x, u2.y);
    const CompConst cc = s.comp[ci];
    const double ri = cc.radius;
    // iForce = EffectiveMass*Gravity (src/MForces.f90:190)
    d3 F = scl3(cc.eff_mass, mk3(s.grav[0], s.grav[1], s.grav[2]));
    d3 T = mk3(0.0, 0.0, 0.0);
    const int cell_p = a.rc_in[i];
    const int st_p = a.st_in[i];
    const int id_p = a.id[i];
    const int cnt_old = a.ccnt_in[i];
    unsigned long long visited = 0;
    int cnt_new = 0;
    const int nn = a.nnb[i];
    for (int k = 0; k < nn; ++k) {
      const int j = a.nbr[(size_t)k * N + i];
      const double4 Q = a.pos_in[j];
      const int cj = (int)Q.w;
      const d3 rij = min_image3(s, mk3(Q.x - P.x, Q.y - P.y, Q.z - P.z));
      const double dsq = dot3(rij, rij);
      const PairConst& pc = s.pair[ci + cj * s.nc];
      const double delta = pc.mindijsq - dsq;
      if (!(delta > 0.0)) continue;  // overlap test (src/MForces.f90:223-224)
      // contact history of the pair (VerifyPairContact, src/MContacts.f90:329)
      int c = -1;
      for (int m = 0; m < cnt_old; ++m)
        if (a.cpart_in[(size_t)m * N + i] == j) { c = m; break; }
      const int id_q = a.id[j];
      int role;
      if (s.hist_mode == 0) {
        role = ref_role(s, cell_p, a.rc_in[j], st_p, a.st_in[j], id_p, id_q);
        if (role < 0) { err |= DEM_KERR_ROLE; role = id_p < id_q; }
      } else {
        role = id_p < id_q;
      }
      const double sg = role ? 1.0 : -1.0;
      ContactIn in;
      in.contacting = c >= 0;
      if (in.contacting) {
        visited |= 1ull << c;
        const double* hp = a.chist_in + (size_t)c * 6 * N + i;
        in.ft0 = scl3(sg, mk3(hp[0], hp[N], hp[2 * N]));
        in.vt0 = scl3(sg, mk3(hp[3 * N], hp[4 * N], hp[5 * N]));
      } else {
        in.ft0 = mk3(0.0, 0.0, 0.0);
        in.vt0 = in.ft0;
        if (role) ++n_col;
      }
      if (role) ++n_pair;
      const CompConst cq = s.comp[cj];
      const double rj = cq.radius;
      const double2* vwq = reinterpret_cast<const double2*>(a.vw_in + 6 * (size_t)j);
      const double2 q0 = vwq[0], q1 = vwq[1], q2 = vwq[2];
      const d3 vj = mk3(q0.x, q0.y, q1.x);
      const d3 wj = mk3(q1.y, q2.x, q2.y);
      const double dij = sqrt(dsq);
      in.nij = div3(rij, dij);
      in.disp_n = scl3(dij - pc.mindij, in.nij);
      in.vij = sub3(vj, vi);
      in.vijc = sub3(in.vij, cross3(add3(scl3(ri, wi), scl3(rj, wj)), in.nij));
      in.delta_dsq = delta;
      in.half_dt = a.half_dt;
      const ContactOut o = common_calc(in, rij, pc.kn, pc.yn, pc.kt, pc.yt, pc.musq);
      F = add3(add3(F, o.fn), o.ft);
      // torque of "i" is (Ri*Nij) x FT, of "j" Rj*(Nij x FT) (src/MForces.f90:259,385)
      if (role) T = add3(T, cross3(scl3(ri, in.nij), o.ft));
      else T = add3(T, scl3(ri, cross3(in.nij, o.ft)));
      if (UPDATE) {
        if (cnt_new < s.max_c) {
          a.cpart_out[(size_t)cnt_new * N + i] = j;
          double* hq = a.chist_out + (size_t)cnt_new * 6 * N + i;
          const d3 sft = scl3(sg, o.ft), svt = scl3(sg, o.vt);
          hq[0] = sft.x; hq[N] = sft.y; hq[2 * N] = sft.z;
          hq[3 * N] = svt.x; hq[4 * N] = svt.y; hq[5 * N] = svt.z;
        } else {
          err |= DEM_KERR_CONTACT_OVERFLOW;
        }
        ++cnt_new;
      }
    }
    if (UPDATE) {
      // Entries not refreshed above belong to non-overlapping pairs.  The
      // reference removes them only when its cell loop visits the pair
      // (RemovePairContact, src/MForces.f90:268); otherwise they persist.
      for (int m = 0; m < cnt_old; ++m) {
        if (visited & (1ull << m)) continue;
        const int q = a.cpart_in[(size_t)m * N + i];
        bool keep = (s.hist_mode == 0) && !ref_adjacent(s, cell_p, a.rc_in[q]);
        if (!keep) continue;
        if (cnt_new < s.max_c) {
          a.cpart_out[(size_t)cnt_new * N + i] = q;
          const double* hp = a.chist_in + (size_t)m * 6 * N + i;
          double* hq = a.chist_out + (size_t)cnt_new * 6 * N + i;
          for (int d = 0; d < 6; ++d) hq[d * N] = hp[d * N];
        } else {
          err |= DEM_KERR_CONTACT_OVERFLOW;
        }
        ++cnt_new;
      }
      if (cnt_new > s.max_c) cnt_new = s.max_c;
      a.ccnt_out[i] = cnt_new;
      n_stored = cnt_new;
      cnt_hw = cnt_new;
    }
    // walls (src/MForces.f90:275-314)
    if (s.nw > 0) {
      unsigned wf = a.wflag[i];
      unsigned wf_new = wf;
      for (int jw = 0; jw < s.nw; ++jw) {
        const WallGeom& w = s.walls[jw];
        const WallConst& wc = s.wallc[ci + jw * s.nc];
        ContactIn in;
        in.contacting = (wf >> jw) & 1u;
        double* hp = a.whist + (size_t)jw * 6 * N + i;
        if (in.contacting) {
          in.ft0 = mk3(hp[0], hp[N], hp[2 * N]);
          in.vt0 = mk3(hp[3 * N], hp[4 * N], hp[5 * N]);
        } else {
          in.ft0 = mk3(0.0, 0.0, 0.0);
          in.vt0 = in.ft0;
        }
        const WallHit wh = wall_overlap(s, w, xi, ri, wc.max_alpha_u, wc.min_aw_sq, wc.max_aw_sq);
        const double dij = wh.dij;
        const d3 nij = wh.nij, qji = wh.qji;
        if (wh.ov < 0) err |= DEM_KERR_CONE_AXIS;
        if (wh.ov > 0) {
          ++n_wall;
          in.nij = nij;
          in.delta_dsq = ri * ri - dij * dij;
          in.disp_n = scl3(dij - ri, nij);
          in.vij = sub3(add3(vget(w.vel), cross3(vget(w.omg), qji)), vi);
          in.vijc = sub3(in.vij, cross3(scl3(ri, wi), nij));
          in.half_dt = a.half_dt;
          const d3 rc = sub3(add3(vget(w.pos), qji), xi);
          const ContactOut o = common_calc(in, rc, wc.knw, wc.ynw, wc.ktw, wc.ytw, wc.muwsq);
          F = add3(add3(F, o.fn), o.ft);
          T = add3(T, cross3(scl3(ri, nij), o.ft));
          if (UPDATE) {
            hp[0] = o.ft.x; hp[N] = o.ft.y; hp[2 * N] = o.ft.z;
            hp[3 * N] = o.vt.x; hp[4 * N] = o.vt.y; hp[5 * N] = o.vt.z;
            wf_new |= 1u << jw;
          }
        } else if (UPDATE && in.contacting) {
          wf_new &= ~(1u << jw);
        }
      }
      if (UPDATE) a.wflag[i] = wf_new;
    }
    // accelerations (src/MForces.f90:330-333)
    const d3 acc = scl3(cc.inv_mass, F);
    const d3 alp = scl3(cc.inv_inertia, T);
    // Particle%Acceleration = Force*InvMass is Force(i,:)*InvMass: same product.
    if (MODE == MODE_ACC) {
      double2* o = reinterpret_cast<double2*>(a.acc_out + 6 * (size_t)i);
      o[0] = make_double2(acc.x, acc.y);
      o[1] = make_double2(acc.z, alp.x);
      o[2] = make_double2(alp.y, alp.z);
    } else if (MODE == MODE_EULER) {
      // Euler (src/MDEM.f90:240-252)
      d3 x = add3(add3(xi, scl3(a.h, vi)), scl3(a.hsq2, acc));
      const d3 v = add3(vi, scl3(a.h, acc));
      const d3 wv = add3(wi, scl3(a.h, alp));
      a.pos_out[i] = make_double4(x.x, x.y, x.z, P.w);
      double2* o = reinterpret_cast<double2*>(a.vw_out + 6 * (size_t)i);
      o[0] = make_double2(v.x, v.y);
      o[1] = make_double2(v.z, wv.x);
      o[2] = make_double2(wv.y, wv.z);
      double2* q = reinterpret_cast<double2*>(a.acc_out + 6 * (size_t)i);
      q[0] = make_double2(acc.x, acc.y);
      q[1] = make_double2(acc.z, alp.x);
      q[2] = make_double2(alp.y, alp.z);
      const int rc = ref_cell(s, x.x, x.y, x.z);
      a.rc_out[i] = rc;
      a.st_out[i] = (rc != cell_p) ? a.stamp_next : st_p;
      const double4 R = a.xref[i];
      const d3 dd = min_image3(s, mk3(x.x - R.x, x.y - R.y, x.z - R.z));
      if (dot3(dd, dd) > s.skin_half_sq) need_rb = 1;
      if (!(x.x == x.x && x.y == x.y && x.z == x.z)) err |= DEM_KERR_NAN;
    } else if (MODE == MODE_PRED) {
      // AB predictor (src/MDEM.f90:297-331)
      d3 sv = mk3(0, 0, 0), sa = sv, sl = sv;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k >= s.nprev) break;
        const double* r = a.slot_rd[k] + i;
        const double c = a.coef[k];
        sv = add3(sv, scl3(c, mk3(r[0], r[N], r[2 * N])));
        sa = add3(sa, scl3(c, mk3(r[3 * N], r[4 * N], r[5 * N])));
        sl = add3(sl, scl3(c, mk3(r[6 * N], r[7 * N], r[8 * N])));
      }
      d3 xp = add3(add3(xi, scl3(a.coef0, vi)), sv);
      const d3 vp = add3(add3(vi, scl3(a.coef0, acc)), sa);
      const d3 wp = add3(add3(wi, scl3(a.coef0, alp)), sl);
      if (s.pbc[0]) xp.x = xp.x - s.lbox[0] * fnint(xp.x * s.invl[0]);
      if (s.pbc[1]) xp.y = xp.y - s.lbox[1] * fnint(xp.y * s.invl[1]);
      if (s.pbc[2]) xp.z = xp.z - s.lbox[2] * fnint(xp.z * s.invl[2]);
      a.pos_out[i] = make_double4(xp.x, xp.y, xp.z, P.w);
      double2* o = reinterpret_cast<double2*>(a.vw_out + 6 * (size_t)i);
      o[0] = make_double2(vp.x, vp.y);
      o[1] = make_double2(vp.z, wp.x);
      o[2] = make_double2(wp.y, wp.z);
      if (s.nprev > 0) {
        double* r = a.slot_wr + i;
        r[0] = vi.x; r[N] = vi.y; r[2 * N] = vi.z;
        r[3 * N] = acc.x; r[4 * N] = acc.y; r[5 * N] = acc.z;
        r[6 * N] = alp.x; r[7 * N] = alp.y; r[8 * N] = alp.z;
      } else {
        double2* q = reinterpret_cast<double2*>(a.acc_out + 6 * (size_t)i);
        q[0] = make_double2(acc.x, acc.y);
        q[1] = make_double2(acc.z, alp.x);
        q[2] = make_double2(alp.y, alp.z);
      }
      const int rc = ref_cell(s, xp.x, xp.y, xp.z);
      a.rc_out[i] = rc;
      a.st_out[i] = (rc != cell_p) ? a.stamp_next : st_p;
      if (a.check_pred) {
        const double4 R = a.xref[i];
        const d3 dd = min_image3(s, mk3(xp.x - R.x, xp.y - R.y, xp.z - R.z));
        if (dot3(dd, dd) > s.skin_half_sq) err |= DEM_KERR_SKIN;
      }
      if (!(xp.x == xp.x && xp.y == xp.y && xp.z == xp.z)) err |= DEM_KERR_NAN;
    } else {  // MODE_CORR
      // AM corrector (src/MDEM.f90:341-367); vi, wi hold the predicted
      // velocities, acc/alp the predicted accelerations.
      if (a.acc_out) {
        double2* q = reinterpret_cast<double2*>(a.acc_out + 6 * (size_t)i);
        q[0] = make_double2(acc.x, acc.y);
        q[1] = make_double2(acc.z, alp.x);
        q[2] = make_double2(alp.y, alp.z);
      }
      const double4 X = a.pos[i];
      const double2* vp0 = reinterpret_cast<const double2*>(a.vw + 6 * (size_t)i);
      const double2 p0 = vp0[0], p1 = vp0[1], p2 = vp0[2];
      const d3 v0 = mk3(p0.x, p0.y, p1.x);
      const d3 w0 = mk3(p1.y, p2.x, p2.y);
      d3 sv = mk3(0, 0, 0), sa = sv, sl = sv, sb = sv;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k >= s.nprev) break;
        const double* r = a.slot_rd[k] + i;
        const double c = a.coef[k];
        const d3 rv = mk3(r[0], r[N], r[2 * N]);
        sv = add3(sv, scl3(c, rv));
        sa = add3(sa, scl3(c, mk3(r[3 * N], r[4 * N], r[5 * N])));
        sl = add3(sl, scl3(c, mk3(r[6 * N], r[7 * N], r[8 * N])));
        sb = add3(sb, scl3(a.ab[k], rv));
      }
      d3 x = add3(add3(mk3(X.x, X.y, X.z), scl3(a.coef0, vi)), sv);
      const d3 v = add3(add3(v0, scl3(a.coef0, acc)), sa);
      const d3 wv = add3(add3(w0, scl3(a.coef0, alp)), sl);
      if (s.pbc[0]) x.x = x.x - s.lbox[0] * fnint(x.x * s.invl[0]);
      if (s.pbc[1]) x.y = x.y - s.lbox[1] * fnint(x.y * s.invl[1]);
      if (s.pbc[2]) x.z = x.z - s.lbox[2] * fnint(x.z * s.invl[2]);
      a.pos[i] = make_double4(x.x, x.y, x.z, X.w);
      double2* o = reinterpret_cast<double2*>(a.vw + 6 * (size_t)i);
      o[0] = make_double2(v.x, v.y);
      o[1] = make_double2(v.z, wv.x);
      o[2] = make_double2(wv.y, wv.z);
      const int rc = ref_cell(s, x.x, x.y, x.z);
      a.rc_out[i] = rc;
      a.st_out[i] = (rc != cell_p) ? a.stamp_next : st_p;
      // escape scan of UpdateParticleList (src/MDEM.f90:624-625)
      if ((x.x > s.phalf[0]) || (x.x < -s.phalf[0]) || (x.y > s.phalf[1]) ||
          (x.y < -s.phalf[1]) || (x.z > s.phalf[2]) || (x.z < -s.phalf[2]))
        n_esc = 1;
      // look-ahead: the next predictor (same expression and order as PRED)
      d3 xn = add3(add3(x, scl3(a.ab0, v)), sb);
      if (s.pbc[0]) xn.x = xn.x - s.lbox[0] * fnint(xn.x * s.invl[0]);
      if (s.pbc[1]) xn.y = xn.y - s.lbox[1] * fnint(xn.y * s.invl[1]);
      if (s.pbc[2]) xn.z = xn.z - s.lbox[2] * fnint(xn.z * s.invl[2]);
      const double4 R = a.xref[i];
      const d3 d1 = min_image3(s, mk3(x.x - R.x, x.y - R.y, x.z - R.z));
      const d3 d2 = min_image3(s, mk3(xn.x - R.x, xn.y - R.y, xn.z - R.z));
      const d3 d3_ = min_image3(s, mk3(xn.x - x.x, xn.y - x.y, xn.z - x.z));
      if (dot3(d1, d1) > s.skin_half_sq || dot3(d2, d2) > s.skin_half_sq) need_rb = 1;
      if (dot3(d3_, d3_) > s.skin_half_sq) err |= DEM_KERR_SKIN;
      if (!(x.x == x.x && x.y == x.y && x.z == x.z)) err |= DEM_KERR_NAN;
    }
  }
  // block reductions -> one atomic per wavefront
  n_col = wave_sum_u64(n_col);
  n_pair = wave_sum_u64(n_pair);
  n_wall = wave_sum_u64(n_wall);
  n_stored = wave_sum_u64(n_stored);
  err = wave_or(err);
  need_rb = wave_or(need_rb);
  n_esc = (int)wave_sum_u64((unsigned long long)n_esc);
  cnt_hw = wave_max(cnt_hw);
  if ((threadIdx.x & 63) == 0) {
    if (n_col) atomicAdd(&ctl->ncol, n_col);
    if (n_pair) atomicAdd(&ctl->pairs[a.par], n_pair);
    if (n_wall) atomicAdd(&ctl->wallc[a.par], n_wall);
    if (UPDATE && n_stored) atomicAdd(&ctl->stored[a.par], n_stored);
    if (err) atomicOr(&ctl->error, err);
    if (need_rb) atomicOr(a.rb_set, 1);
    if (UPDATE && cnt_hw) atomicMax(&ctl->max_cnt, cnt_hw);
    if (MODE == MODE_CORR && n_esc) {
      atomicAdd(&ctl->escaped, n_esc);
      atomicMin(&ctl->stop_step, a.step_idx);
    }
  }
}

// ---------------------------------------------------------------------------
// Reference-grid cell assignment of uploaded positions (CurrentCell3D and the
// transfer bookkeeping of RedistributeParticles3D, src/MCells.f90:607-643).
__global__ void k_refcell(SysConst s, const double4* pos, const int* rc_prev, const int* st_prev,
                          int* rc_out, int* st_out, int stamp_next, int fresh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.np) return;
  const double4 P = pos[i];
  const int rc = ref_cell(s, P.x, P.y, P.z);
  rc_out[i] = rc;
  st_out[i] = fresh ? st_prev[i] : ((rc != rc_prev[i]) ? stamp_next : st_prev[i]);
}

// Displacement check of uploaded positions against the neighbour-list origin.
__global__ void k_check_disp(SysConst s, const double4* pos, const double4* xref, int* rb_set) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int need = 0;
  if (i < s.np) {
    const double4 P = pos[i], R = xref[i];
    const d3 dd = min_image3(s, mk3(P.x - R.x, P.y - R.y, P.z - R.z));
    need = dot3(dd, dd) > s.skin_half_sq;
  }
  need = wave_or(need);
  if ((threadIdx.x & 63) == 0 && need) atomicOr(rb_set, 1);
}

// ---------------------------------------------------------------------------
// Counting sort of particles into the neighbour-search grid.  All kernels
// take a gate: when *gate == 0 (and force == 0) they return at once, so the
// rebuild can be enqueued every step without a host round trip.

__device__ __forceinline__ bool gate_closed(const int* gate, int force, const Ctl* ctl) {
  if (*(volatile const int*)&ctl->escaped) return true;
  return !force && !*(volatile const int*)gate;
}

__global__ void k_fill_i32(int* p, int n, int v, const int* gate, int force, const Ctl* ctl) {
  if (gate_closed(gate, force, ctl)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_bin_count(SysConst s, const double4* pos, int* key, int* count, const int* gate,
                            int force, const Ctl* ctl) {
  if (gate_closed(gate, force, ctl)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.np; i += gridDim.x * blockDim.x) {
    const double4 P = pos[i];
    const int c = grid_cell(s, P.x, P.y, P.z);
    key[i] = c;
    atomicAdd(&count[c], 1);
  }
}

#define SCAN_ITEMS 8
#define SCAN_TILE (DEM_BLOCK * SCAN_ITEMS)

// Inclusive scan of one value per lane across the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_scan_incl(int v) {
  const int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread across the block; returns the block
// total in *total.
__device__ __forceinline__ int block_scan_excl(int v, int* total) {
  __shared__ int wsum[DEM_BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int inc = wave_scan_incl(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < DEM_BLOCK / 64; ++w) {
    if (w < wid) off += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ __launch_bounds__(DEM_BLOCK) void k_scan_tiles(const int* in, int n, int* tile_sum,
                                                          const int* gate, int force,
                                                          const Ctl* ctl) {
  if (gate_closed(gate, force, ctl)) return;
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) s += (base + k < n) ? in[base + k] : 0;
  int tot;
  block_scan_excl(s, &tot);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(DEM_BLOCK) void k_scan_spine(int* tile_sum, int ntiles,
                                                          const int* gate, int force,
                                                          const Ctl* ctl) {
  if (gate_closed(gate, force, ctl)) return;
  int carry = 0;
  for (int b = 0; b < ntiles; b += DEM_BLOCK) {
    const int t = b + threadIdx.x;
    const int v = t < ntiles ? tile_sum[t] : 0;
    int tot;
    const int ex = block_scan_excl(v, &tot);
    if (t < ntiles) tile_sum[t] = carry + ex;
    carry += tot;
  }
}

__global__ __launch_bounds__(DEM_BLOCK) void k_scan_apply(const int* in, int n,
                                                          const int* tile_off, int* start,
                                                          int* cursor, const int* gate,
                                                          int force, const Ctl* ctl) {
  if (gate_closed(gate, force, ctl)) return;
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int tot;
  int ex = block_scan_excl(s, &tot) + tile_off[blockIdx.x];
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) {
      start[base + k] = ex;
      cursor[base + k] = ex;
    }
    ex += v[k];
  }
  if (base <= n - 1 && n - 1 < base + SCAN_ITEMS) start[n] = ex;
}

__global__ void k_bin_scatter(int np, const int* key, int* cursor, int* cell_idx,
                              const int* gate, int force, const Ctl* ctl) {
  if (gate_closed(gate, force, ctl)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
    const int p = atomicAdd(&cursor[key[i]], 1);
    cell_idx[p] = i;
  }
}

// Sort each cell's slice by particle index (insertion sort; cells hold a few
// particles), so the order does not depend on atomic arrival order.
__global__ void k_cell_sort(int ncell, const int* start, int* cell_idx, const int* gate,
                            int force, const Ctl* ctl) {
  if (gate_closed(gate, force, ctl)) return;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
    const int b = start[c], e = start[c + 1];
    for (int p = b + 1; p < e; ++p) {
      const int v = cell_idx[p];
      int q = p - 1;
      while (q >= b && cell_idx[q] > v) {
        cell_idx[q + 1] = cell_idx[q];
        --q;
      }
      cell_idx[q + 1] = v;
    }
  }
}

// Build the directed neighbour list: all j != i with |r_ij| < r_i + r_j + skin
// (minimum image), scanning the 27 cells around i's cell in a fixed order.
__global__ __launch_bounds__(DEM_BLOCK) void k_nbr_build(SysConst s, const double4* pos,
                                                         const int* key, const int* start,
                                                         const int* cell_idx, int* nbr, int* nnb,
                                                         double4* xref, const int* gate,
                                                         int force, Ctl* ctl) {
  if (gate_closed(gate, force, ctl)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->rebuilds, 1ull);
  const size_t N = (size_t)s.np;
  const int i = xcd_block() * DEM_BLOCK + threadIdx.x;
  int cnt = 0, err = 0;
  if (i < s.np) {
    const double4 P = pos[i];
    const int ci = (int)P.w;
    const int k = key[i];
    const int gx = k % s.gn[0], gy = (k / s.gn[0]) % s.gn[1], gz = k / (s.gn[0] * s.gn[1]);
    for (int dz = -1; dz <= 1; ++dz) {
      int cz = gz + dz;
      if (s.pbc[2]) {
        if (s.gn[2] == 1 && dz != 0) continue;
        if (s.gn[2] == 2 && dz == 1) continue;
        cz = (cz + s.gn[2]) % s.gn[2];
      } else if (cz < 0 || cz >= s.gn[2]) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        int cy = gy + dy;
        if (s.pbc[1]) {
          if (s.gn[1] == 1 && dy != 0) continue;
          if (s.gn[1] == 2 && dy == 1) continue;
          cy = (cy + s.gn[1]) % s.gn[1];
        } else if (cy < 0 || cy >= s.gn[1]) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          int cx = gx + dx;
          if (s.pbc[0]) {
            if (s.gn[0] == 1 && dx != 0) continue;
            if (s.gn[0] == 2 && dx == 1) continue;
            cx = (cx + s.gn[0]) % s.gn[0];
          } else if (cx < 0 || cx >= s.gn[0]) continue;
          const int c = cx + (cy + cz * s.gn[1]) * s.gn[0];
          const int e = start[c + 1];
          for (int p = start[c]; p < e; ++p) {
            const int j = cell_idx[p];
            if (j == i) continue;
            const double4 Q = pos[j];
            const d3 r = min_image3(s, mk3(Q.x - P.x, Q.y - P.y, Q.z - P.z));
            const double dsq = dot3(r, r);
            if (dsq < s.pair[ci + (int)Q.w * s.nc].rc2) {
              if (cnt < s.max_nb) nbr[(size_t)cnt * N + i] = j;
              else err = DEM_KERR_NBR_OVERFLOW;
              ++cnt;
            }
          }
        }
      }
    }
    nnb[i] = cnt < s.max_nb ? cnt : s.max_nb;
    xref[i] = P;
  }
  unsigned long long tot = wave_sum_u64((unsigned long long)(cnt < s.max_nb ? cnt : s.max_nb));
  err = wave_or(err);
  const int hw = wave_max(cnt);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&ctl->nbpairs, tot);
    if (err) atomicOr(&ctl->error, err);
    atomicMax(&ctl->max_nnb, hw);
  }
}

// ---------------------------------------------------------------------------
// Resort: move every per-particle array from old order to new order
// (new i <- old cell_idx[i]); contact partners are renumbered.

__global__ void k_old_to_new(int np, const int* cell_idx, int* old_to_new) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) old_to_new[cell_idx[i]] = i;
}

struct PermArgs {
  const int* src;       // new -> old
  const int* old_to_new;
  const double4* pos;  double4* pos2;
  const double* vw;    double* vw2;
  const double* acc;   double* acc2;
  const double* ring;  double* ring2;  int nring;  // 9*nprev components
  const int* id;       int* id2;
  const int* rc;       int* rc2;
  const int* st;       int* st2;
  const int* ccnt;     int* ccnt2;
  const int* cpart;    int* cpart2;
  const double* chist; double* chist2;
  const unsigned* wflag; unsigned* wflag2;
  const double* whist;   double* whist2;
  int nw;
};

__global__ void k_permute(SysConst s, PermArgs a) {
  const size_t N = (size_t)s.np;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.np) return;
  const int o = a.src[i];
  a.pos2[i] = a.pos[o];
  for (int d = 0; d < 6; ++d) a.vw2[6 * (size_t)i + d] = a.vw[6 * (size_t)o + d];
  for (int d = 0; d < 6; ++d) a.acc2[6 * (size_t)i + d] = a.acc[6 * (size_t)o + d];
  for (int c = 0; c < a.nring; ++c) a.ring2[c * N + i] = a.ring[c * N + o];
  a.id2[i] = a.id[o];
  a.rc2[i] = a.rc[o];
  a.st2[i] = a.st[o];
  const int n = a.ccnt[o];
  a.ccnt2[i] = n;
  for (int m = 0; m < n; ++m) {
    a.cpart2[m * N + i] = a.old_to_new[a.cpart[m * N + o]];
    for (int d = 0; d < 6; ++d) a.chist2[(m * 6 + d) * N + i] = a.chist[(m * 6 + d) * N + o];
  }
  if (a.nw > 0) {
    a.wflag2[i] = a.wflag[o];
    for (int c = 0; c < 6 * a.nw; ++c) a.whist2[c * N + i] = a.whist[c * N + o];
  }
}

// ---------------------------------------------------------------------------
// Host <-> device reordering.

// Particle records (TParticle layout) in reference order -> internal arrays.
__global__ void k_aos_in(int np, const dem_particle_t* p, const int* id, double4* pos, double* vw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const dem_particle_t& r = p[id[i]];
  pos[i] = make_double4(r.position[0], r.position[1], r.position[2], (double)(r.index - 1));
  double* o = vw + 6 * (size_t)i;
  o[0] = r.velocity[0]; o[1] = r.velocity[1]; o[2] = r.velocity[2];
  o[3] = r.omega[0]; o[4] = r.omega[1]; o[5] = r.omega[2];
}

__global__ void k_aos_out(int np, dem_particle_t* p, const int* id, const double* acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  dem_particle_t& r = p[id[i]];
  const double* a = acc + 6 * (size_t)i;
  r.acceleration[0] = a[0]; r.acceleration[1] = a[1]; r.acceleration[2] = a[2];
  r.alpha[0] = a[3]; r.alpha[1] = a[4]; r.alpha[2] = a[5];
}

// Gathers internal arrays into reference-order SoA columns:
// out[c*np + id[i]] = in component c of particle i.
__global__ void k_to_ref_d(int np, int ncomp, const double* in, int stride_rec, int stride_comp,
                           const int* id, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int r = id[i];
  for (int c = 0; c < ncomp; ++c)
    out[(size_t)c * np + r] = in[(size_t)i * stride_rec + (size_t)c * stride_comp];
}

__global__ void k_to_ref_i(int np, const int* in, const int* id, int* out, int add) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  out[id[i]] = in[i] + add;
}

// Deterministic partial sums for the kinetic energy (src/MDEM.f90:199-213):
// each block sums its particles in index order, the host sums the blocks.
__global__ __launch_bounds__(DEM_BLOCK) void k_energy(SysConst s, const double4* pos,
                                                      const double* vw, double* part) {
  __shared__ double buf[DEM_BLOCK];
  const int i = blockIdx.x * DEM_BLOCK + threadIdx.x;
  double e = 0.0;
  if (i < s.np) {
    const CompConst c = s.comp[(int)pos[i].w];
    const double* v = vw + 6 * (size_t)i;
    const double v2 = dot3(mk3(v[0], v[1], v[2]), mk3(v[0], v[1], v[2]));
    const double w2 = dot3(mk3(v[3], v[4], v[5]), mk3(v[3], v[4], v[5]));
    e = c.mass * v2 + c.inertia * w2;
  }
  buf[threadIdx.x] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < DEM_BLOCK; ++k) t += buf[k];
    part[blockIdx.x] = t;
  }
}
