// kernels.hpp -- gfx950 kernels of the DEM per-timestep path.
//
// Layout in HBM (internal particle order = neighbour-grid cell order of the
// last resort).  Two state buffers b = 0, 1 hold Particle and Predicted
// (src/MDEM.f90:20); which one is Particle is host state:
//   pos[b]  double4[N]  x, y, z, component (0-based, as a double)
//   kin[b]  Kin[N]      64-byte line: velocity, angular velocity, reference id,
//                       and the reference-grid cell state (CellIndex and the
//                       cell-list entry stamp) of the positions in pos[b]
//   ring    double[nprev][9][N]  ABM ring Previous(:,i) (src/MDEM.f90:14-24)
//   edges (two sets, selected by the device word *epar; a rebuild writes the
//         other set and flips it), sliced ELL with 64-particle slices (one
//         slice per wavefront): the k-th neighbour of particle i is edge
//         e = soff[i/64] + 64*k + i%64, so a wavefront reading the k-th
//         edge of its 64 particles touches one contiguous run.
//         cnt  int[N]          neighbours of each particle (low 16 bits) and
//                              how many of them come first as contacts
//         soff int[N/64+1]     first edge of each slice (64 * slice width)
//         nbr  int[E]          neighbour j of edge e
//         flag u8[E]           bit 0: the pair has stored history
//         hist double2[3][E]   planes (FTx,FTy) (FTz,VTx) (VTy,VTz): tangential
//                              force FT and velocity VT of the pair as the
//                              reference stores them (TPartner%Vector1/Vector2),
//                              FP64 in both precision modes
//   wall history: wflag uint[N] (bit w: contact with wall w), whist [nw][6][N]
//   stale pairs (StaleTab): stored pairs that have left the neighbour list
//         while their reference cells are not adjacent, by sorted 64-bit
//         pair key, with their history (reference semantics, below)
//
// A force pass runs one lane per particle.  Each lane walks its own edges in
// list order (all lanes of a wavefront read their k-th edge together),
// gathers the partner, evaluates the contact and accumulates force and
// torque in registers: no atomics and no inter-lane traffic, so the sums are
// bitwise reproducible.  The same lane then applies the integrator to its
// particle.
#pragma once
#pragma clang fp contract(off)

#include "physics.hpp"
#include "mixed.hpp"

#ifndef DEM_PASS_WAVES
#define DEM_PASS_WAVES 3
#endif
// Counter slots of a pass: block b adds its counts into slot b % DEM_RED_SLOTS,
// and the reduction after the pass reads and clears min(blocks, DEM_RED_SLOTS)
// slots, so a pass leaves them zero for the next one.
#define DEM_RED_SLOTS 256
// Default occupancy target of the force passes (wavefronts per SIMD): the
// register budget is 512 / DEM_PASS_WAVES VGPRs.

enum { MODE_ACC = 0, MODE_EULER = 1, MODE_PRED = 2, MODE_CORR = 3 };

struct __align__(64) Kin {
  double2 v0, v1, v2;  // vx vy | vz wx | wy wz
  int id;              // reference id, 0-based
  int rc;              // reference-grid cell (0-based) of the positions
  int st;              // entry stamp into that cell's list
  int rxyz;            // packed coordinates of rc (when SysConst.rc_pack)
};

__device__ __forceinline__ Kin load_kin(const Kin* k) {
  const int4* q = reinterpret_cast<const int4*>(k);
  int4 a = q[0], b = q[1], c = q[2], d = q[3];
  Kin r;
  r.v0 = make_double2(__hiloint2double(a.y, a.x), __hiloint2double(a.w, a.z));
  r.v1 = make_double2(__hiloint2double(b.y, b.x), __hiloint2double(b.w, b.z));
  r.v2 = make_double2(__hiloint2double(c.y, c.x), __hiloint2double(c.w, c.z));
  r.id = d.x;
  r.rc = d.y;
  r.st = d.z;
  r.rxyz = d.w;
  return r;
}

__device__ __forceinline__ void store_kin(Kin* k, d3 v, d3 w, int id, RefCell rc, int st) {
  double2* q = reinterpret_cast<double2*>(k);
  q[0] = make_double2(v.x, v.y);
  q[1] = make_double2(v.z, w.x);
  q[2] = make_double2(w.y, w.z);
  reinterpret_cast<int4*>(k)[3] = make_int4(id, rc.lin, st, rc.xyz);
}

__device__ __forceinline__ d3 kin_v(const Kin& k) { return mk3(k.v0.x, k.v0.y, k.v1.x); }
__device__ __forceinline__ d3 kin_w(const Kin& k) { return mk3(k.v1.y, k.v2.x, k.v2.y); }

// Blocks are dispatched round-robin over the 8 XCDs; map consecutive block
// ranges to one XCD so a spatial slab of particles shares one L2.  (The
// force passes choose between this and the dispatch order: PassArgs::xcd_map.)
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
// Inclusive scan of one value per lane across the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_scan_incl(int v) {
  const int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Stale pairs.  The reference forgets a stored pair only when the cell loop
// of an updating ComputeAccelerations visits the pair (reference cells the
// same or neighbours) and finds no overlap (RemovePairContact,
// src/MForces.f90:262-270, src/MContacts.f90:427-437).  A pair that leaves
// the neighbour list while its reference cells are not adjacent therefore
// keeps its history there, and a later contact starts from it.  The table
// holds such pairs between rebuilds:
//   key   u64[DEM_STALE_ROWS]   (id_a + 1) << 32 | (id_b + 1), id_a < id_b;
//                               rows [0, cnt[0]) sorted ascending
//   slot  int2[DEM_STALE_ROWS]  the two particles' indices (current order)
//   hist  double2[3][DEM_STALE_ROWS]  the history as the edges hold it
//   state u8[DEM_STALE_ROWS]    0 kept; 1 restored into the list by the last
//                               rebuild; 2 forgotten (visited, no overlap)
//   cnt   int[3]                rows; rows appended by the running rebuild;
//                               pairs lost because the table was full
// A rebuild looks every new edge without carried history up in the table
// (binary search) and restores a kept row; it appends the stored pairs its
// old list loses while their reference cells are not adjacent (rows
// cnt[0] + k); then the flip kernel merges: kept rows and new rows, sorted
// by key, the others dropped.  The reduction after an updating pass marks
// the rows whose reference cells are adjacent at that pass as forgotten (the
// pair is no neighbour, so it does not overlap).
#define DEM_STALE_ROWS 4096
struct StaleTab {
  unsigned long long* key;
  int2* slot;
  double2* hist;
  unsigned char* state;
  int* cnt;
};

// Tiles of the staged force passes: particles [DEM_TILE*b, DEM_TILE*b +
// DEM_TILE) form tile b (consecutive in cell-key order, so a compact region).
// Its halo is the set of neighbours outside it, ascending, at most
// DEM_HALO_CAP of them; the slot of an edge is the index of its neighbour in
// the tile's staged records: j - DEM_TILE*b inside the tile, DEM_TILE + the
// halo row outside it, DEM_SLOT_NONE when the halo overflowed.
#define DEM_TILE DEM_BLOCK
#define DEM_HALO_CAP 512
#define DEM_SLOT_NONE 0xffff

struct EdgeSets {
  int* cnt[2];
  int* soff[2];
  int* nbr[2];
  unsigned char* flag[2];
  double2* hist[2];  // three planes of ecap entries each
  unsigned short* slot[2];  // staged slot of each edge's neighbour
  int* halo[2];             // DEM_HALO_CAP rows per tile
  int* hcnt[2];             // halo rows of each tile
  long long ecap;
  int* epar;  // device word: index of the current set
  StaleTab st;  // null pointers in the canonical history mode
  int halo_on;  // the rebuilds list tile halos and slots (a staged shape is in use)
  int halo_rows;  // rows listed per tile, <= DEM_HALO_CAP (option "halo_rows")
};

__host__ __device__ __forceinline__ unsigned long long pair_key(int ida, int idb) {
  const int lo = ida < idb ? ida : idb, hi = ida < idb ? idb : ida;
  return ((unsigned long long)(lo + 1) << 32) | (unsigned long long)(hi + 1);
}

// Row of key in the sorted rows [0, n) of the table, or -1.
__device__ __forceinline__ int stale_find(const unsigned long long* key, int n,
                                          unsigned long long k) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && key[lo] == k) ? lo : -1;
}

// cnt[i] holds the list length of particle i in its low 16 bits and, once
// the list is filled, the number of front (contact-first) entries above them
// (nbr_fill_body).
__host__ __device__ __forceinline__ int cnt_deg(int c) { return c & 0xffff; }
__host__ __device__ __forceinline__ int cnt_front(int c) { return c >> 16; }

__device__ __forceinline__ long long sell_edge(const int* soff, int i, int k) {
  return (long long)soff[i >> 6] + 64LL * k + (i & 63);
}

struct PassArgs {
  // evaluated state (Particle or Predicted): positions, kin with cell state
  const double4* pos_in;
  const Kin* kin_in;
  // Particle state (CORR reads and rewrites it in place)
  double4* pos;
  Kin* kin;
  // PRED: Predicted; EULER: the next Particle buffer
  double4* pos_out;
  Kin* kin_out;
  double2* acc_out;     // ACC/EULER/PRED(order 1): a, alpha; CORR: predicted (optional)
  const double* slot_rd[3];  // ring slots in ipc order (k = 1..nprev)
  double* slot_wr;           // PRED: slot that becomes ipc(1)
  double coef0, coef[3];     // AB0/AB(k) (PRED), AM0/AM(k) (CORR)
  double ab0, ab[3];         // CORR: AB coefficients for the look-ahead predictor
  double h, hsq2;            // EULER sub-step and 0.5*h**2
  double half_dt;            // HalfTimeStep
  int stamp_next;            // entry stamp of the next evaluation
  int check_pred;            // PRED: verify the predicted positions against xref
  int step_idx;              // CORR: step index within the current dem_step call
  int par;                   // evaluation parity (counters)
  const double4* xref;
  EdgeSets E;
  unsigned* wflag;
  double* whist;
  int* rb_clear;             // rebuild flag consumed before this evaluation
  int* rb_set;               // rebuild flag for the next evaluation
  unsigned* part;            // counter slots (DEM_RED_SLOTS): collisions, pairs, walls, stored
  Ctl* ctl;
  // PRED: Predicted%Index by reference id when it differs from Particle%Index
  // (after removals, until the next resize, src/MDEM.f90:636-640); else null
  const double* pred_w;
  // fluid drag of every particle of this evaluation (k_fluid_drag), or null
  const double* drag;
  // Sums of the AB predictor over the ring (9 planes of np: velocity,
  // acceleration, angular acceleration terms).  The corrector of a step reads
  // the ring slots the next predictor reads, so it forms the next predictor's
  // sums in the same order and stores them (psum_wr); that predictor then
  // reads 72 bytes per particle instead of the three 72-byte slots
  // (psum_rd; null = read the slots).
  const double* psum_rd;
  double* psum_wr;
  // block order: 1 = consecutive particle ranges per XCD (xcd_block), 0 =
  // the dispatch order (consecutive blocks on different XCDs, all XCDs in
  // one region of the particles at a time)
  int xcd_map;
  int pad_;
};

// The pair history and the ABM ring slots and sums are streamed: read or
// written once per pass, by one lane.  Their loads and stores carry the
// non-temporal hint, so the caches favour the partner records, which every
// neighbour of a particle reads again.  (The same hint on the neighbour
// indices and flags, or on the corrector's own Particle record, measured
// slower: the next pass reads those again.)
typedef double dem_dv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_nt(const double2* p) {
  const dem_dv2 v = __builtin_nontemporal_load(reinterpret_cast<const dem_dv2*>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
  dem_dv2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<dem_dv2*>(p));
}
template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}
// Nine planar components r[0], r[N], ..., r[8N] as three vectors.
__device__ __forceinline__ void ld_ring(const double* r, size_t N, d3& v, d3& a, d3& l) {
  v = mk3(ld_nt(r), ld_nt(r + N), ld_nt(r + 2 * N));
  a = mk3(ld_nt(r + 3 * N), ld_nt(r + 4 * N), ld_nt(r + 5 * N));
  l = mk3(ld_nt(r + 6 * N), ld_nt(r + 7 * N), ld_nt(r + 8 * N));
}
__device__ __forceinline__ void st_ring(double* r, size_t N, d3 v, d3 a, d3 l) {
  st_nt(r, v.x); st_nt(r + N, v.y); st_nt(r + 2 * N, v.z);
  st_nt(r + 3 * N, a.x); st_nt(r + 4 * N, a.y); st_nt(r + 5 * N, a.z);
  st_nt(r + 6 * N, l.x); st_nt(r + 7 * N, l.y); st_nt(r + 8 * N, l.z);
}

__device__ __forceinline__ void store_d3x2(double2* o, d3 a, d3 b) {
  o[0] = make_double2(a.x, a.y);
  o[1] = make_double2(a.z, b.x);
  o[2] = make_double2(b.y, b.z);
}

// Component tables of up to DEM_LDS_NC components are staged in LDS by every
// force pass, so the per-edge lookups of the pair constants (indexed by the
// partner's component) cost an LDS read instead of a cache round trip.
#define DEM_LDS_NC 16

// Pair contacts of one particle, in list order (src/MForces.f90:203-271).
//
// One lane walks its particle's list.  The list rebuild puts the pairs that
// touch (or carry history) first, nfront of them, and the rest after them
// (nbr_fill_body), so the partner records an edge needs are known from its
// place in the list:
//  - a front edge is a contact almost always: the partner's position, its
//    64-byte Kin line (velocities, cell state) and the stored history are
//    requested together, and the edge waits for one memory round trip;
//  - a back edge is a contact rarely (a pair that came into contact since
//    the rebuild): only the partner's position (and history, if any) is
//    requested, and the Kin line follows only when the pair overlaps.
// With FRONT = 0 every edge is treated as a front edge (the autotuner tries
// both).  The index and flag of edge k+1 are requested before the records
// of edge k.  The arithmetic and its order do not depend on FRONT.
struct EdgeAcc {
  d3 F, T;          // FP64 sums (and gravity)
  KahanF3 Ff, Tf;   // compensated FP32 sums of the mixed mode
  unsigned long long n_col, n_pair, n_stored;
  int err;
};

// The stored history of edge e as the three (FTx,FTy) (FTz,VTx) (VTy,VTz)
// pairs.
__device__ __forceinline__ void load_hist(const double2* __restrict__ hist, long long ecap,
                                          long long e, double2& h0, double2& h1, double2& h2) {
  h0 = ld_nt(hist + e);
  h1 = ld_nt(hist + ecap + e);
  h2 = ld_nt(hist + 2 * ecap + e);
}

// The contribution of one contact to its particle's sums: fn and ft (FP64;
// the FP32 sums of the mixed mode round them there) and the torque (the
// mixed mode's FP32 torque, widened exactly).
struct EdgeOut {
  d3 fn, ft, tq;
};

// The contact of an edge whose pair overlaps (delta = MinDijSq - DijSq > 0):
// the role of the pair ("i" of the reference), CommonCalculations, the
// torque and, in an updating pass, the new history (src/MForces.f90:223-263).
// KJ is the partner's Kin line.  Counters and errors go to acc; the force
// and torque contributions are returned (acc_add sums them).
template <bool UPDATE, bool MIXED>
__device__ __forceinline__ EdgeOut edge_eval(const SysConst& s, const PassArgs& a,
                                             const PairConst& pc, double rj, const d3& rij,
                                             double dsq, double delta, const Kin& KJ,
                                             unsigned char fl, double2 h0, double2 h1, double2 h2,
                                             unsigned char* __restrict__ flag,
                                             double2* __restrict__ hist, long long ecap,
                                             long long e, d3 vi, d3 wi, double ri, int cell_p,
                                             int xyz_p, int st_p, int id_p, EdgeAcc& acc) {
  int role;
  if (s.hist_mode == 0) {
    role = ref_role(s, cell_p, xyz_p, KJ.rc, KJ.rxyz, st_p, KJ.st, id_p, KJ.id);
    if (role < 0) { acc.err |= DEM_KERR_ROLE; role = id_p < KJ.id; }
  } else {
    role = id_p < KJ.id;
  }
  const double sg = role ? 1.0 : -1.0;
  ContactIn in;
  in.contacting = fl & 1;
  in.ft0 = scl3(sg, mk3(h0.x, h0.y, h1.x));
  in.vt0 = scl3(sg, mk3(h1.y, h2.x, h2.y));
  if (!in.contacting && role) ++acc.n_col;
  if (role) ++acc.n_pair;
  const double dij = sqrt(dsq);
  in.nij = div3(rij, dij);
  in.disp_n = scl3(dij - pc.mindij, in.nij);
  in.vij = sub3(kin_v(KJ), vi);
  in.vijc = sub3(in.vij, cross3(add3(scl3(ri, wi), scl3(rj, kin_w(KJ))), in.nij));
  in.delta_dsq = delta;
  in.half_dt = a.half_dt;
  const ContactOut o = common_calc(in, rij, pc.kn, pc.yn, pc.kt, pc.yt, pc.musq);
  EdgeOut r;
  r.fn = o.fn;
  r.ft = o.ft;
  if (MIXED) {
    // the contact in FP64, its torque formed in FP32 as the sums (mixed.hpp)
    const f3 t = sclf((float)ri, crossf(tof(in.nij), tof(o.ft)));
    r.tq = mk3((double)t.x, (double)t.y, (double)t.z);
  } else {
    // torque of "i" is (Ri*Nij) x FT, of "j" Rj*(Nij x FT) (src/MForces.f90:259,385)
    r.tq = role ? cross3(scl3(ri, in.nij), o.ft) : scl3(ri, cross3(in.nij, o.ft));
  }
  if (UPDATE) {
    const d3 f = scl3(sg, o.ft), g = scl3(sg, o.vt);
    st_nt(hist + e, make_double2(f.x, f.y));
    st_nt(hist + ecap + e, make_double2(f.z, g.x));
    st_nt(hist + 2 * ecap + e, make_double2(g.y, g.z));
    if (!(fl & 1)) flag[e] = 1;
    ++acc.n_stored;
  }
  return r;
}

// iForce = iForce + ijForceN + ijForceT, iTorque = iTorque + torque
// (src/MForces.f90:381-385), in FP64, or in the compensated FP32 sums of the
// mixed mode (mixed.hpp).
template <bool MIXED>
__device__ __forceinline__ void acc_add(EdgeAcc& acc, const d3& fn, const d3& ft, const d3& tq) {
  if (MIXED) {
    kahan_add3(acc.Ff, tof(fn));
    kahan_add3(acc.Ff, tof(ft));
    kahan_add3(acc.Tf, tof(tq));
  } else {
    acc.F = add3(add3(acc.F, fn), ft);
    acc.T = add3(acc.T, tq);
  }
}

template <bool UPDATE, bool MIXED>
__device__ __forceinline__ void edge_contact(const SysConst& s, const PassArgs& a,
                                             const PairConst& pc, double rj, const d3& rij,
                                             double dsq, double delta, const Kin& KJ,
                                             unsigned char fl, double2 h0, double2 h1, double2 h2,
                                             unsigned char* __restrict__ flag,
                                             double2* __restrict__ hist, long long ecap,
                                             long long e, d3 vi, d3 wi, double ri, int cell_p,
                                             int xyz_p, int st_p, int id_p, EdgeAcc& acc) {
  const EdgeOut o = edge_eval<UPDATE, MIXED>(s, a, pc, rj, rij, dsq, delta, KJ, fl, h0, h1, h2,
                                             flag, hist, ecap, e, vi, wi, ri, cell_p, xyz_p,
                                             st_p, id_p, acc);
  acc_add<MIXED>(acc, o.fn, o.ft, o.tq);
}

// A stored pair that no longer overlaps, in an updating pass.  The reference
// removes it only when its cell loop visits it (RemovePairContact,
// src/MForces.f90:268): reference cells the same or neighbours (rc, rxyz: the
// partner's cell state); otherwise it persists.
__device__ __forceinline__ void edge_separated(const SysConst& s, int cell_p, int xyz_p, int rc,
                                               int rxyz, unsigned char* __restrict__ flag,
                                               long long e, EdgeAcc& acc) {
  const bool keep = (s.hist_mode == 0) && !ref_adjacent(s, cell_p, xyz_p, rc, rxyz);
  if (keep) ++acc.n_stored;
  else flag[e] = 0;
}

template <bool UPDATE, bool MIXED, int FRONT>
__device__ __forceinline__ void pair_loop(const SysConst& s, const PassArgs& a,
                                          const PairConst* ptab, const CompConst* ctab,
                                          const int* __restrict__ nbr,
                                          unsigned char* __restrict__ flag,
                                          double2* __restrict__ hist, long long ecap,
                                          long long e, int deg, int nfront, const double4& P,
                                          d3 vi, d3 wi, int ci, double ri, int cell_p, int xyz_p,
                                          int st_p, int id_p, EdgeAcc& acc) {
  if (!FRONT) nfront = deg;
  int j_n = 0;
  unsigned char fl_n = 0;
  if (deg > 0) {
    j_n = nbr[e];
    fl_n = flag[e];
  }
  for (int k = 0; k < deg; ++k, e += 64) {
    const int j = j_n;
    const unsigned char fl = fl_n;
    if (k + 1 < deg) {
      j_n = nbr[e + 64];
      fl_n = flag[e + 64];
    }
    const bool front = k < nfront;
    const double4 Q = a.pos_in[j];
    Kin KJ;
    if (front) KJ = load_kin(a.kin_in + j);
    double2 h0 = make_double2(0.0, 0.0), h1 = h0, h2 = h0;
    if (fl & 1) load_hist(hist, ecap, e, h0, h1, h2);
    const int cj = (int)Q.w;
    const d3 rij = min_image3(s, mk3(Q.x - P.x, Q.y - P.y, Q.z - P.z));
    const double dsq = dot3(rij, rij);
    const PairConst& pc = ptab[ci + cj * s.nc];
    const double delta = pc.mindijsq - dsq;   // overlap test (src/MForces.f90:223-224)
    if (delta > 0.0) {
      if (!front) KJ = load_kin(a.kin_in + j);
      edge_contact<UPDATE, MIXED>(s, a, pc, ctab[cj].radius, rij, dsq, delta, KJ, fl, h0, h1, h2,
                                  flag, hist, ecap, e, vi, wi, ri, cell_p, xyz_p, st_p, id_p, acc);
    } else if (UPDATE && (fl & 1)) {
      int rc, rxyz;
      if (front) {
        rc = KJ.rc;
        rxyz = KJ.rxyz;
      } else {
        const int4 c = reinterpret_cast<const int4*>(a.kin_in + j)[3];
        rc = c.y;
        rxyz = c.w;
      }
      edge_separated(s, cell_p, xyz_p, rc, rxyz, flag, e, acc);
    }
  }
}

// The same walk with the partner records staged in LDS by the pass (tile
// records and the tile's halo, pass_body_v): the position of every partner,
// and the Kin line of partners inside the tile, come from LDS; the Kin line of
// a halo partner is gathered from memory (its index from the staged halo
// rows), and an edge whose partner is not staged (DEM_SLOT_NONE) gathers both.
// The arithmetic and its order are those of pair_loop.
template <bool UPDATE, bool MIXED, int FRONT>
__device__ __forceinline__ void pair_loop_staged(
    const SysConst& s, const PassArgs& a, const PairConst* ptab, const CompConst* ctab,
    const unsigned short* __restrict__ slot, const int* __restrict__ nbr,
    unsigned char* __restrict__ flag, double2* __restrict__ hist, long long ecap, long long e,
    int deg, int nfront, const double4& P, d3 vi, d3 wi, int ci, double ri, int cell_p, int xyz_p,
    int st_p, int id_p, const double4* sp, const Kin* sk, const int* shalo, EdgeAcc& acc) {
  if (!FRONT) nfront = deg;
  unsigned sl_n = 0;
  unsigned char fl_n = 0;
  if (deg > 0) {
    sl_n = slot[e];
    fl_n = flag[e];
  }
  for (int k = 0; k < deg; ++k, e += 64) {
    const unsigned sl = sl_n;
    const unsigned char fl = fl_n;
    if (k + 1 < deg) {
      sl_n = slot[e + 64];
      fl_n = flag[e + 64];
    }
    const bool front = k < nfront;
    const bool inside = sl < DEM_TILE;
    int j = 0;
    double4 Q;
    Kin KJ;
    if (inside) {
      Q = sp[sl];
      if (front) KJ = sk[sl];
    } else {
      if (sl != DEM_SLOT_NONE) {
        Q = sp[sl];
        j = shalo[sl - DEM_TILE];
      } else {
        j = nbr[e];
        Q = a.pos_in[j];
      }
      if (front) KJ = load_kin(a.kin_in + j);
    }
    double2 h0 = make_double2(0.0, 0.0), h1 = h0, h2 = h0;
    if (fl & 1) load_hist(hist, ecap, e, h0, h1, h2);
    const int cj = (int)Q.w;
    const d3 rij = min_image3(s, mk3(Q.x - P.x, Q.y - P.y, Q.z - P.z));
    const double dsq = dot3(rij, rij);
    const PairConst& pc = ptab[ci + cj * s.nc];
    const double delta = pc.mindijsq - dsq;   // overlap test (src/MForces.f90:223-224)
    if (delta > 0.0) {
      if (!front) KJ = inside ? sk[sl] : load_kin(a.kin_in + j);
      edge_contact<UPDATE, MIXED>(s, a, pc, ctab[cj].radius, rij, dsq, delta, KJ, fl, h0, h1, h2,
                                  flag, hist, ecap, e, vi, wi, ri, cell_p, xyz_p, st_p, id_p, acc);
    } else if (UPDATE && (fl & 1)) {
      int rc, rxyz;
      if (front) {
        rc = KJ.rc;
        rxyz = KJ.rxyz;
      } else if (inside) {
        rc = sk[sl].rc;
        rxyz = sk[sl].rxyz;
      } else {
        const int4 c = reinterpret_cast<const int4*>(a.kin_in + j)[3];
        rc = c.y;
        rxyz = c.w;
      }
      edge_separated(s, cell_p, xyz_p, rc, rxyz, flag, e, acc);
    }
  }
}
// One force pass over the particles [blk*BS, blk*BS+BS) of one system,
// BS lanes per block.  Neither the block size nor FRONT changes the
// arithmetic of a contact, the order of a particle's sums or the integer
// counter sums, so every launch shape gives the same results bit for bit.
// The arguments of a pass that change from step to step (prepare_step in
// batch.hpp, step_args on the host); the rest of PassArgs stays fixed.
struct StepVars {
  int stamp_next, check_pred, step_idx, par;
  int* rb_clear;
  int* rb_set;
  const double* slot_rd[3];
  double* slot_wr;
  const double* psum_rd;
  double* psum_wr;
};
__device__ __forceinline__ StepVars step_vars(const PassArgs& a) {
  StepVars v;
  v.stamp_next = a.stamp_next;
  v.check_pred = a.check_pred;
  v.step_idx = a.step_idx;
  v.par = a.par;
  v.rb_clear = a.rb_clear;
  v.rb_set = a.rb_set;
  for (int k = 0; k < 3; ++k) v.slot_rd[k] = a.slot_rd[k];
  v.slot_wr = a.slot_wr;
  v.psum_rd = a.psum_rd;
  v.psum_wr = a.psum_wr;
  return v;
}

// Pass variants (the FRONT template argument of the passes): bit 0 is the
// front rule of the pair walk (1: the partner's Kin line is requested with
// its position only for front edges); bit 1 stages the partner records of
// the block's tile in LDS (pair_loop_staged; BS == DEM_TILE), which leaves
// room in LDS for the pair tables of at most DEM_LDS_NC_STAGED components.
#define DEM_PV_STAGED 2
#define DEM_LDS_NC_STAGED 8

template <int MODE, bool UPDATE, bool MIXED, int BS = DEM_BLOCK, int FRONT = 1>
__device__ __forceinline__ void pass_body_v(const SysConst& s, const PassArgs& a,
                                            const StepVars& stv, int blk) {
  constexpr bool STAGED = (FRONT & DEM_PV_STAGED) != 0;
  constexpr int NCL = STAGED ? DEM_LDS_NC_STAGED : DEM_LDS_NC;
  static_assert(!STAGED || BS == DEM_TILE, "a staged pass runs one block per tile");

  __shared__ PairConst sh_pair[NCL * NCL];
  __shared__ CompConst sh_comp[NCL];
  // staged partner records: the tile's own particles, then its halo
  __shared__ double4 st_pos[STAGED ? DEM_TILE + DEM_HALO_CAP : 1];
  __shared__ Kin st_kin[STAGED ? DEM_TILE : 1];
  __shared__ int st_halo[STAGED ? DEM_HALO_CAP : 1];
  Ctl* ctl = a.ctl;
  // Escapes are published between passes (k_pass_reduce_*), so every block of
  // a pass reads the same value here.
  if (ctl->escaped) return;
  const int t = threadIdx.x;
  if (blk == 0 && t == 0) *stv.rb_clear = 0;
  const bool lds_tab = s.nc <= NCL;
  if (lds_tab) {
    for (int q = t; q < s.nc * s.nc; q += BS) sh_pair[q] = s.pair[q];
    if (t < s.nc) sh_comp[t] = s.comp[t];
  }
  const size_t N = (size_t)s.np;
  const int ep = *a.E.epar;

  const int i = blk * BS + t;
  const bool active = i < s.np;
  unsigned long long n_wall = 0;
  int need_rb = 0, n_esc = 0;

  double4 P = make_double4(0, 0, 0, 0);
  d3 vi = mk3(0, 0, 0), wi = vi;
  int cell_p = 0, xyz_p = 0, st_p = 0, id_p = 0, deg = 0, nfront = 0;
  long long e = 0;
  if (active) {
    P = a.pos_in[i];
    const Kin K = load_kin(a.kin_in + i);
    vi = kin_v(K);
    wi = kin_w(K);
    cell_p = K.rc;
    xyz_p = K.rxyz;
    st_p = K.st;
    id_p = K.id;
    const int cn = a.E.cnt[ep][i];
    deg = cnt_deg(cn);
    nfront = cnt_front(cn);
    e = sell_edge(a.E.soff[ep], i, 0);
    if (STAGED) {
      st_pos[t] = P;
      st_kin[t] = K;
    }
  }
  if (STAGED) {
    // the halo's positions (independent gathers, all in flight together)
    const int hn = a.E.hcnt[ep][blk];
    const int* hrow = a.E.halo[ep] + (size_t)blk * DEM_HALO_CAP;
    for (int k = t; k < hn; k += BS) {
      const int j = hrow[k];
      st_halo[k] = j;
      st_pos[DEM_TILE + k] = a.pos_in[j];
    }
  }
  const int ci = (int)P.w;
  const CompConst cc = s.comp[active ? ci : 0];
  const double ri = cc.radius;
  EdgeAcc acc;
  // iForce = EffectiveMass*Gravity (src/MForces.f90:190)
  acc.F = scl3(cc.eff_mass, mk3(s.grav[0], s.grav[1], s.grav[2]));
  acc.T = mk3(0.0, 0.0, 0.0);
  acc.Ff.s = acc.Ff.c = acc.Tf.s = acc.Tf.c = mkf(0.0f, 0.0f, 0.0f);
  acc.n_col = acc.n_pair = acc.n_stored = 0;
  acc.err = 0;
  if (STAGED || lds_tab) __syncthreads();
  if (STAGED)
    pair_loop_staged<UPDATE, MIXED, FRONT & 1>(
        s, a, lds_tab ? sh_pair : s.pair, lds_tab ? sh_comp : s.comp, a.E.slot[ep], a.E.nbr[ep],
        a.E.flag[ep], a.E.hist[ep], a.E.ecap, e, deg, nfront, P, vi, wi, ci, ri, cell_p, xyz_p,
        st_p, id_p, st_pos, st_kin, st_halo, acc);
  else
    pair_loop<UPDATE, MIXED, FRONT & 1>(s, a, lds_tab ? sh_pair : s.pair,
                                        lds_tab ? sh_comp : s.comp, a.E.nbr[ep], a.E.flag[ep],
                                        a.E.hist[ep], a.E.ecap, e, deg, nfront, P, vi, wi, ci, ri,
                                        cell_p, xyz_p, st_p, id_p, acc);
  d3 F = acc.F, T = acc.T;
  if (MIXED) {
    F = add3(F, kahan_value(acc.Ff));
    T = add3(T, kahan_value(acc.Tf));
  }
  unsigned long long n_col = acc.n_col, n_pair = acc.n_pair, n_stored = acc.n_stored;
  int err = acc.err;

  if (active) {
    const d3 xi = mk3(P.x, P.y, P.z);
    // walls (src/MForces.f90:275-314)
    if (s.nw > 0) {
      const unsigned wf = a.wflag[i];
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
        if (wh.ov < 0) err |= DEM_KERR_CONE_AXIS;
        if (wh.ov > 0) {
          ++n_wall;
          in.nij = wh.nij;
          in.delta_dsq = ri * ri - wh.dij * wh.dij;
          in.disp_n = scl3(wh.dij - ri, wh.nij);
          in.vij = sub3(add3(vget(w.vel), cross3(vget(w.omg), wh.qji)), vi);
          in.vijc = sub3(in.vij, cross3(scl3(ri, wi), wh.nij));
          in.half_dt = a.half_dt;
          const d3 rc = sub3(add3(vget(w.pos), wh.qji), xi);
          const ContactOut o = common_calc(in, rc, wc.knw, wc.ynw, wc.ktw, wc.ytw, wc.muwsq);
          F = add3(add3(F, o.fn), o.ft);
          T = add3(T, cross3(scl3(ri, wh.nij), o.ft));
          if (UPDATE) {
            hp[0] = o.ft.x; hp[N] = o.ft.y; hp[2 * N] = o.ft.z;
            hp[3 * N] = o.vt.x; hp[4 * N] = o.vt.y; hp[5 * N] = o.vt.z;
            wf_new |= 1u << jw;
          }
        } else if (UPDATE && in.contacting) {
          wf_new &= ~(1u << jw);
        }
      }
      if (UPDATE && wf_new != wf) a.wflag[i] = wf_new;
    }
    // iForce = iForce + DragForce(iPart) (src/MForces.f90:316)
    if (a.drag) F = add3(F, vget(a.drag + 3 * (size_t)i));
    // accelerations (src/MForces.f90:330-333)
    const d3 acc = scl3(cc.inv_mass, F);
    const d3 alp = scl3(cc.inv_inertia, T);
    if (MODE == MODE_ACC) {
      store_d3x2(a.acc_out + 3 * (size_t)i, acc, alp);
    } else if (MODE == MODE_EULER) {
      // Euler (src/MDEM.f90:240-252)
      const d3 x = add3(add3(xi, scl3(a.h, vi)), scl3(a.hsq2, acc));
      const d3 v = add3(vi, scl3(a.h, acc));
      const d3 wv = add3(wi, scl3(a.h, alp));
      a.pos_out[i] = make_double4(x.x, x.y, x.z, P.w);
      const RefCell rc = ref_cell(s, x.x, x.y, x.z);
      store_kin(a.kin_out + i, v, wv, id_p, rc, (rc.lin != cell_p) ? stv.stamp_next : st_p);
      store_d3x2(a.acc_out + 3 * (size_t)i, acc, alp);
      const double4 R = a.xref[i];
      const d3 dd = min_image3(s, mk3(x.x - R.x, x.y - R.y, x.z - R.z));
      if (dot3(dd, dd) > s.skin_half_sq) need_rb = 1;
      if (!(x.x == x.x && x.y == x.y && x.z == x.z)) err |= DEM_KERR_NAN;
    } else if (MODE == MODE_PRED) {
      // AB predictor (src/MDEM.f90:297-331)
      d3 sv = mk3(0, 0, 0), sa = sv, sl = sv;
      if (stv.psum_rd) {
        ld_ring(stv.psum_rd + i, N, sv, sa, sl);
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (k >= s.nprev) break;
          const double c = a.coef[k];
          d3 rv, ra, rl;
          ld_ring(stv.slot_rd[k] + i, N, rv, ra, rl);
          sv = add3(sv, scl3(c, rv));
          sa = add3(sa, scl3(c, ra));
          sl = add3(sl, scl3(c, rl));
        }
      }
      d3 xp = add3(add3(xi, scl3(a.coef0, vi)), sv);
      const d3 vp = add3(add3(vi, scl3(a.coef0, acc)), sa);
      const d3 wp = add3(add3(wi, scl3(a.coef0, alp)), sl);
      if (s.pbc[0]) xp.x = xp.x - s.lbox[0] * fnint(xp.x * s.invl[0]);
      if (s.pbc[1]) xp.y = xp.y - s.lbox[1] * fnint(xp.y * s.invl[1]);
      if (s.pbc[2]) xp.z = xp.z - s.lbox[2] * fnint(xp.z * s.invl[2]);
      a.pos_out[i] = make_double4(xp.x, xp.y, xp.z, a.pred_w ? a.pred_w[id_p] : P.w);
      const RefCell rc = ref_cell(s, xp.x, xp.y, xp.z);
      store_kin(a.kin_out + i, vp, wp, id_p, rc, (rc.lin != cell_p) ? stv.stamp_next : st_p);
      if (s.nprev > 0) {
        st_ring(stv.slot_wr + i, N, vi, acc, alp);
      } else {
        store_d3x2(a.acc_out + 3 * (size_t)i, acc, alp);
      }
      if (stv.check_pred) {
        const double4 R = a.xref[i];
        const d3 dd = min_image3(s, mk3(xp.x - R.x, xp.y - R.y, xp.z - R.z));
        if (dot3(dd, dd) > s.skin_half_sq) err |= DEM_KERR_SKIN;
      }
      if (!(xp.x == xp.x && xp.y == xp.y && xp.z == xp.z)) err |= DEM_KERR_NAN;
    } else {  // MODE_CORR
      // AM corrector (src/MDEM.f90:341-367); vi, wi hold the predicted
      // velocities, acc/alp the predicted accelerations.
      if (a.acc_out) store_d3x2(a.acc_out + 3 * (size_t)i, acc, alp);
      const double4 X = a.pos[i];
      const Kin K0 = load_kin(a.kin + i);
      const d3 v0 = kin_v(K0), w0 = kin_w(K0);
      // sb, sba, sbl: the next predictor's sums (same slots, same order)
      d3 sv = mk3(0, 0, 0), sa = sv, sl = sv, sb = sv, sba = sv, sbl = sv;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k >= s.nprev) break;
        const double c = a.coef[k];
        d3 rv, ra, rl;
        ld_ring(stv.slot_rd[k] + i, N, rv, ra, rl);
        sv = add3(sv, scl3(c, rv));
        sa = add3(sa, scl3(c, ra));
        sl = add3(sl, scl3(c, rl));
        sb = add3(sb, scl3(a.ab[k], rv));
        sba = add3(sba, scl3(a.ab[k], ra));
        sbl = add3(sbl, scl3(a.ab[k], rl));
      }
      if (stv.psum_wr) {
        st_ring(stv.psum_wr + i, N, sb, sba, sbl);
      }
      d3 x = add3(add3(mk3(X.x, X.y, X.z), scl3(a.coef0, vi)), sv);
      const d3 v = add3(add3(v0, scl3(a.coef0, acc)), sa);
      const d3 wv = add3(add3(w0, scl3(a.coef0, alp)), sl);
      if (s.pbc[0]) x.x = x.x - s.lbox[0] * fnint(x.x * s.invl[0]);
      if (s.pbc[1]) x.y = x.y - s.lbox[1] * fnint(x.y * s.invl[1]);
      if (s.pbc[2]) x.z = x.z - s.lbox[2] * fnint(x.z * s.invl[2]);
      a.pos[i] = make_double4(x.x, x.y, x.z, X.w);
      const RefCell rc = ref_cell(s, x.x, x.y, x.z);
      store_kin(a.kin + i, v, wv, id_p, rc, (rc.lin != cell_p) ? stv.stamp_next : st_p);
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
  // Counters: block sums to a per-block slot (k_pass_reduce_* add them up);
  // flags, which are rare, by atomics.
  __shared__ unsigned sh_cnt[BS / 64][4];
  const unsigned c0 = (unsigned)wave_sum_u64(n_col), c1 = (unsigned)wave_sum_u64(n_pair);
  const unsigned c2 = (unsigned)wave_sum_u64(n_wall), c3 = (unsigned)wave_sum_u64(n_stored);
  err = wave_or(err);
  need_rb = wave_or(need_rb);
  n_esc = (int)wave_sum_u64((unsigned long long)n_esc);
  if ((t & 63) == 0) {
    sh_cnt[t >> 6][0] = c0;
    sh_cnt[t >> 6][1] = c1;
    sh_cnt[t >> 6][2] = c2;
    sh_cnt[t >> 6][3] = c3;
    if (err) atomicOr(&ctl->error, err);
    if (need_rb) atomicOr(stv.rb_set, 1);
    if (MODE == MODE_CORR && n_esc) {
      atomicAdd(&ctl->esc_pending, n_esc);
      atomicMin(&ctl->stop_step, stv.step_idx);
    }
  }
  __syncthreads();
  if (t < 4) {
    unsigned v = 0;
    for (int w = 0; w < BS / 64; ++w) v += sh_cnt[w][t];
    // integer adds: the totals do not depend on the order of the blocks
    if (v) atomicAdd(&a.part[4 * (size_t)(blk % DEM_RED_SLOTS) + t], v);
  }
}

template <int MODE, bool UPDATE, bool MIXED, int BS = DEM_BLOCK, int FRONT = 1>
__device__ __forceinline__ void pass_body(const SysConst& s, const PassArgs& a, int blk) {
  pass_body_v<MODE, UPDATE, MIXED, BS, FRONT>(s, a, step_vars(a), blk);
}

// Launch shapes of the force passes: lanes per block, the occupancy the
// register allocation targets (wavefronts per SIMD, 512/WAVES VGPRs each;
// above 3 the pass spills to scratch) and the pass variant FRONT (bit 0, the
// front rule of pair_loop: 1 = the partner's Kin line is gathered with its
// position only for the front edges, 0 = for every edge; bit 1: partner
// records staged in LDS per tile, pair_loop_staged).  Shape 0 is the
// default; the others are candidates of the autotuner (dem_autotune).  Every
// shape gives the same results bit for bit.
#ifndef DEM_NSHAPES
#define DEM_NSHAPES 8
#endif
template <int S>
struct PassShape;
#define DEM_PASS_SHAPE(S, B, W, F)           \
  template <>                                \
  struct PassShape<S> {                      \
    static constexpr int block = B;          \
    static constexpr int waves = W;          \
    static constexpr int front = F;          \
  };
DEM_PASS_SHAPE(0, 256, DEM_PASS_WAVES, 1)
DEM_PASS_SHAPE(1, 128, DEM_PASS_WAVES, 1)
DEM_PASS_SHAPE(2, 64, DEM_PASS_WAVES, 1)
DEM_PASS_SHAPE(3, 256, 2, 1)
DEM_PASS_SHAPE(4, 256, DEM_PASS_WAVES, 1 | DEM_PV_STAGED)
DEM_PASS_SHAPE(5, 256, DEM_PASS_WAVES, 0 | DEM_PV_STAGED)
DEM_PASS_SHAPE(6, 128, DEM_PASS_WAVES, 0)
DEM_PASS_SHAPE(7, 256, DEM_PASS_WAVES, 0)
#undef DEM_PASS_SHAPE
#define DEM_MIN_PASS_BLOCK 64

template <int MODE, bool UPDATE, bool MIXED, int BS = DEM_BLOCK, int WAVES = DEM_PASS_WAVES,
          int FRONT = 1>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_pass(
    SysConst s, PassArgs a) {
  pass_body<MODE, UPDATE, MIXED, BS, FRONT>(s, a, a.xcd_map ? xcd_block() : (int)blockIdx.x);
}

// Sums the per-block counters of the pass just run into the control block
// and publishes the escapes its corrector found.  Runs as one block.
// Sums of the four counters over the per-block slots [b0, b1) by one block
// of 1024 threads; the totals are valid in thread 0.
// Counter slots [b0, b1) summed and cleared by one block of NT threads; the
// totals on thread 0.
template <int NT = 1024>
__device__ __forceinline__ void sum_slots(unsigned* part, int b0, int b1,
                                          unsigned long long s[4]) {
  __shared__ unsigned long long sh[NT / 64][4];
  const int t = threadIdx.x;
  unsigned long long v[4] = {0, 0, 0, 0};
  for (int b = b0 + t; b < b1; b += NT)
    for (int c = 0; c < 4; ++c) {
      v[c] += part[4 * (size_t)b + c];
      part[4 * (size_t)b + c] = 0u;
    }
  for (int c = 0; c < 4; ++c) v[c] = wave_sum_u64(v[c]);
  if ((t & 63) == 0)
    for (int c = 0; c < 4; ++c) sh[t >> 6][c] = v[c];
  __syncthreads();
  for (int c = 0; c < 4; ++c) s[c] = 0;
  if (t == 0)
    for (int w = 0; w < NT / 64; ++w)
      for (int c = 0; c < 4; ++c) s[c] += sh[w][c];
}

// Counter totals s of a pass into the control block, and the escapes its
// corrector found published (thread 0 only).
__device__ __forceinline__ void publish_counts(Ctl* ctl, const unsigned long long s[4], int par,
                                               int update) {
  ctl->ncol += s[0];
  ctl->pairs[par] = s[1];
  ctl->wallc[par] = s[2];
  if (update) {
    ctl->stored[par] = s[3];
    ctl->pairs_upd_sum += s[1];
    ctl->upd_evals += 1;
  }
  if (ctl->esc_pending) {
    ctl->escaped = ctl->esc_pending;
    ctl->esc_pending = 0;
  }
}

// After an updating pass: the stale pairs whose reference cells are now
// adjacent were visited by the reference's cell loop without overlapping
// (they are no neighbours) and are forgotten; returns, on thread 0 of the
// block (NT threads), how many rows are kept.
template <int NT = 1024>
__device__ __forceinline__ unsigned long long stale_visit(const SysConst& s, const Kin* kin,
                                                          const StaleTab& T) {
  __shared__ unsigned long long sh_kept[NT / 64];
  const int t = threadIdx.x;
  const int n = min(T.cnt[0], DEM_STALE_ROWS);
  unsigned long long kept = 0;
  for (int k = t; k < n; k += NT) {
    if (T.state[k] != 0) continue;
    const int2 q = T.slot[k];
    const int4 a = reinterpret_cast<const int4*>(kin + q.x)[3];
    const int4 b = reinterpret_cast<const int4*>(kin + q.y)[3];
    if (ref_adjacent(s, a.y, a.w, b.y, b.w)) T.state[k] = 2;
    else ++kept;
  }
  kept = wave_sum_u64(kept);
  if ((t & 63) == 0) sh_kept[t >> 6] = kept;
  __syncthreads();
  kept = 0;
  if (t == 0)
    for (int w = 0; w < NT / 64; ++w) kept += sh_kept[w];
  return kept;
}

// nslots <= DEM_RED_SLOTS counter slots of a pass, summed by one block of NT
// threads, cleared and published; after an updating pass, the visit of the
// stale pairs, whose kept rows count as stored pairs (both directions).
template <int NT = 1024>
__device__ __forceinline__ void pass_reduce_body(Ctl* ctl, unsigned* part, int nslots, int par,
                                                 int update, const SysConst& sc, const Kin* kin,
                                                 const StaleTab& T) {
  if (ctl->escaped) return;  // the pass did not run
  const int t = threadIdx.x;
  unsigned long long s[4];
  sum_slots<NT>(part, 0, nslots, s);
  if (update && sc.hist_mode == 0 && T.cnt) s[3] += 2 * stale_visit<NT>(sc, kin, T);
  if (t == 0) publish_counts(ctl, s, par, update);
}
__global__ __launch_bounds__(1024) void k_pass_reduce(Ctl* ctl, unsigned* part, int nblk,
                                                      int par, int update, SysConst sc,
                                                      const Kin* kin, StaleTab T) {
  pass_reduce_body(ctl, part, min(nblk, DEM_RED_SLOTS), par, update, sc, kin, T);
}

// MoveWalls(Time) (src/MDEM.f90:392-408): every axis of every wall, one
// thread per wall.  Phase is reduced by 2 pi with truncation toward zero as
// int(Phase/TwoPi) does.  Does nothing once a step had escapes (the passes
// after it do not run either).
__device__ __forceinline__ void move_walls_body(WallGeom* walls, const WallMotion* wm, int nw,
                                                double time, const Ctl* ctl) {
  if (ctl->escaped) return;
  const double two_pi = 2.0 * 3.14159265358979323846;
  for (int j = threadIdx.x; j < nw; j += blockDim.x) {
    const WallMotion m = wm[j];
    for (int d = 0; d < 3; ++d) {
      double ph = m.tpf[d] * time + m.phase[d];
      ph = ph - two_pi * trunc(ph / two_pi);
      walls[j].pos[d] = m.cf[d] + m.amp[d] * cos(ph);
      walls[j].vel[d] = m.vamp[d] * sin(ph);
    }
  }
}
__global__ void k_move_walls(WallGeom* walls, const WallMotion* wm, int nw, double time,
                             const Ctl* ctl) {
  move_walls_body(walls, wm, nw, time, ctl);
}

// ---------------------------------------------------------------------------
// Cell state of positions set from outside (upload, the ComputeAccelerations
// drop-in, the PBC wrap after the Euler start-up): CurrentCell3D plus the
// transfer bookkeeping of RedistributeParticles3D (src/MCells.f90:607-643)
// against the cell state of the previous evaluation (prev; may alias kin).
// wrap != 0 applies ApplyPeriodicBoundaryConditions (src/MDEM.f90:373-388)
// to the positions first; wrap < 0 only assigns the cells and keeps the
// stamps (a fresh upload, whose stamps are the given list ranks).
__global__ void k_refcell(SysConst s, double4* pos, const Kin* prev, Kin* kin, int stamp_next,
                          int wrap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.np) return;
  double4 P = pos[i];
  if (wrap > 0) {
    if (s.pbc[0]) P.x = P.x - s.lbox[0] * fnint(P.x * s.invl[0]);
    if (s.pbc[1]) P.y = P.y - s.lbox[1] * fnint(P.y * s.invl[1]);
    if (s.pbc[2]) P.z = P.z - s.lbox[2] * fnint(P.z * s.invl[2]);
    pos[i] = P;
  }
  const int rc_prev = prev[i].rc, st_prev = prev[i].st;
  const RefCell rc = ref_cell(s, P.x, P.y, P.z);
  kin[i].rc = rc.lin;
  kin[i].rxyz = rc.xyz;
  kin[i].st = (wrap < 0 || rc.lin == rc_prev) ? st_prev : stamp_next;
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
  if (ctl->escaped) return true;
  return !force && !*gate;
}

// Kernel bodies take their block index blk and block count nblk explicitly:
// the single-system kernels pass blockIdx.x / gridDim.x, the batched ones
// (batch.hpp) the block's position within its own system.

__device__ __forceinline__ void bin_count_body(const SysConst& s, const double4* pos, int* key,
                                               int* count, const int* gate, int force,
                                               const Ctl* ctl, int blk, int nblk) {
  if (gate_closed(gate, force, ctl)) return;
  for (int i = blk * DEM_BLOCK + threadIdx.x; i < s.np; i += nblk * DEM_BLOCK) {
    const double4 P = pos[i];
    const int c = grid_cell(s, P.x, P.y, P.z);
    key[i] = c;
    atomicAdd(&count[c], 1);
  }
}
__global__ void k_bin_count(SysConst s, const double4* pos, int* key, int* count, const int* gate,
                            int force, const Ctl* ctl) {
  bin_count_body(s, pos, key, count, gate, force, ctl, blockIdx.x, gridDim.x);
}

#define SCAN_ITEMS 8
#define SCAN_TILE (DEM_BLOCK * SCAN_ITEMS)


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

__device__ __forceinline__ void scan_tiles_body(const int* in, int n, int* tile_sum,
                                                const int* gate, int force, const Ctl* ctl,
                                                int blk) {
  if (gate_closed(gate, force, ctl)) return;
  const int base = blk * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) s += (base + k < n) ? in[base + k] : 0;
  int tot;
  block_scan_excl(s, &tot);
  if (threadIdx.x == 0) tile_sum[blk] = tot;
}
__global__ __launch_bounds__(DEM_BLOCK) void k_scan_tiles(const int* in, int n, int* tile_sum,
                                                          const int* gate, int force,
                                                          const Ctl* ctl) {
  scan_tiles_body(in, n, tile_sum, gate, force, ctl, blockIdx.x);
}

// Exclusive scan of in[0..n) into start[0..n] (start[n] = total); cursor (may
// be null) receives a copy of start[0..n).  With out_sel != null the output
// is the slice offsets of the edge set being built (1 - *epar).  The offset
// of tile blk is the sum of the tile totals before it (scan_tiles_body),
// which the block adds up itself.  clear != 0 sets in[0..n) back to zero
// once read (the cell counts, so the next binning starts from zero).
__device__ __forceinline__ void scan_apply_body(int* in, int n, const int* tile_sum, int* start,
                                                int* cursor, const EdgeSets* out_sel,
                                                int clear, const int* gate, int force,
                                                const Ctl* ctl, int blk) {
  if (gate_closed(gate, force, ctl)) return;
  if (out_sel) start = out_sel->soff[1 - *out_sel->epar];
  int pre = 0;
  for (int q = threadIdx.x; q < blk; q += DEM_BLOCK) pre += tile_sum[q];
  int tile_off;
  block_scan_excl(pre, &tile_off);
  const int base = blk * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  if (clear)
    for (int k = 0; k < SCAN_ITEMS; ++k)
      if (base + k < n) in[base + k] = 0;
  int tot;
  int ex = block_scan_excl(s, &tot) + tile_off;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) {
      start[base + k] = ex;
      if (cursor) cursor[base + k] = ex;
    }
    ex += v[k];
  }
  if (base <= n - 1 && n - 1 < base + SCAN_ITEMS) start[n] = ex;
}
__global__ __launch_bounds__(DEM_BLOCK) void k_scan_apply(int* in, int n, const int* tile_sum,
                                                          int* start, int* cursor,
                                                          const EdgeSets* out_sel, int clear,
                                                          const int* gate, int force,
                                                          const Ctl* ctl) {
  scan_apply_body(in, n, tile_sum, start, cursor, out_sel, clear, gate, force, ctl, blockIdx.x);
}

__device__ __forceinline__ void bin_scatter_body(int np, const int* key, int* cursor,
                                                 int* cell_idx, const int* gate, int force,
                                                 const Ctl* ctl, int blk, int nblk) {
  if (gate_closed(gate, force, ctl)) return;
  for (int i = blk * DEM_BLOCK + threadIdx.x; i < np; i += nblk * DEM_BLOCK) {
    const int p = atomicAdd(&cursor[key[i]], 1);
    cell_idx[p] = i;
  }
}
__global__ void k_bin_scatter(int np, const int* key, int* cursor, int* cell_idx,
                              const int* gate, int force, const Ctl* ctl) {
  bin_scatter_body(np, key, cursor, cell_idx, gate, force, ctl, blockIdx.x, gridDim.x);
}

// Sort each cell's slice by particle index (insertion sort; cells hold a few
// particles), so the order does not depend on atomic arrival order.
__device__ __forceinline__ void cell_sort_body(int ncell, const int* start, int* cell_idx,
                                               const int* gate, int force, const Ctl* ctl,
                                               int blk, int nblk) {
  if (gate_closed(gate, force, ctl)) return;
  for (int c = blk * DEM_BLOCK + threadIdx.x; c < ncell; c += nblk * DEM_BLOCK) {
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
__global__ void k_cell_sort(int ncell, const int* start, int* cell_idx, const int* gate,
                            int force, const Ctl* ctl) {
  cell_sort_body(ncell, start, cell_idx, gate, force, ctl, blockIdx.x, gridDim.x);
}

// Visits every j != i with |r_ij|^2 < (r_i + r_j + skin)^2 (minimum image),
// scanning the 27 cells around i's cell in a fixed order (dz, dy, dx, then
// the cell's slice in order): fn(j, |r_ij|^2, component of j).  rc2 is the
// table of (r_i + r_j + skin)^2 by component pair (LDS, staged by the
// caller).  The slices of a row of three cells are requested together, and
// their particles B at a time (indices, then positions), so a lane waits for
// a few round trips per row instead of two per candidate.
template <int B = 4, typename F>
__device__ __forceinline__ void for_each_neighbour(const SysConst& s, const double4* pos,
                                                   int i, const double4& P,
                                                   const int* start, const int* cell_idx,
                                                   const double* rc2, F&& fn) {
  const int ci = (int)P.w;
  const Int3 g = grid_coords(s, P.x, P.y, P.z);
  const int gx = g.x, gy = g.y, gz = g.z;
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
      int cb[3], ce[3];
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const int dx = u - 1;
        int cx = gx + dx;
        bool skip = false;
        if (s.pbc[0]) {
          if (s.gn[0] == 1 && dx != 0) skip = true;
          if (s.gn[0] == 2 && dx == 1) skip = true;
          cx = (cx + s.gn[0]) % s.gn[0];
        } else if (cx < 0 || cx >= s.gn[0]) {
          skip = true;
        }
        cb[u] = ce[u] = 0;
        if (!skip) {
          const int c = grid_key(s, cx, cy, cz);
          cb[u] = start[c];
          ce[u] = start[c + 1];
        }
      }
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        for (int p0 = cb[u]; p0 < ce[u]; p0 += B) {
          int jj[B];
          double4 QQ[B];
#pragma unroll
          for (int v = 0; v < B; ++v) jj[v] = (p0 + v < ce[u]) ? cell_idx[p0 + v] : -1;
#pragma unroll
          for (int v = 0; v < B; ++v)
            if (jj[v] >= 0) QQ[v] = pos[jj[v]];
#pragma unroll
          for (int v = 0; v < B; ++v) {
            const int j = jj[v];
            if (j < 0 || j == i) continue;
            const double4 Q = QQ[v];
            const d3 r = min_image3(s, mk3(Q.x - P.x, Q.y - P.y, Q.z - P.z));
            const double dsq = dot3(r, r);
            const int cj = (int)Q.w;
            const int q = ci + cj * s.nc;
            if (dsq < (s.cut_all ? s.rc2_all : rc2 ? rc2[q] : s.pair[q].rc2)) fn(j, dsq, cj);
          }
        }
      }
    }
  }
}

// The pair tables the rebuild kernels read per candidate, (r_i + r_j +
// skin)^2 and MinDijSq by component pair, staged in LDS by one block of
// DEM_BLOCK threads (ok = 0: more than DEM_LDS_NC components, read from
// SysConst instead).
struct NbrTab {
  double rc2[DEM_LDS_NC * DEM_LDS_NC];
  double mindsq[DEM_LDS_NC * DEM_LDS_NC];
  bool ok;
};
__device__ __forceinline__ void stage_nbr_tab(const SysConst& s, NbrTab& t) {
  t.ok = s.nc <= DEM_LDS_NC;
  if (t.ok)
    for (int q = threadIdx.x; q < s.nc * s.nc; q += DEM_BLOCK) {
      t.rc2[q] = s.pair[q].rc2;
      t.mindsq[q] = s.pair[q].mindijsq;
    }
  __syncthreads();
}

// Pass 1 of the list build: neighbours per particle (capped at max_nb) into
// cnt of the set being built, and 64 * (widest list) of each 64-particle
// slice into slice_w.
__device__ __forceinline__ void nbr_count_body(const SysConst& s, const double4* pos,
                                               const int* start, const int* cell_idx,
                                               const EdgeSets& E, int* slice_w,
                                               const int* gate, int force, Ctl* ctl, int blk) {
  if (gate_closed(gate, force, ctl)) return;
  __shared__ NbrTab tab;
  stage_nbr_tab(s, tab);
  const int q = 1 - *E.epar;
  const int i = blk * DEM_BLOCK + threadIdx.x;
  int n = 0, err = 0;
  if (i < s.np) {
    const double4 P = pos[i];
    for_each_neighbour<4>(s, pos, i, P, start, cell_idx, tab.ok ? tab.rc2 : nullptr,
                          [&](int, double, int) { ++n; });
    if (n > s.max_nb) {
      err = DEM_KERR_NBR_OVERFLOW;
      n = s.max_nb;
    }
    E.cnt[q][i] = n;
  }
  err = wave_or(err);
  const int hw = wave_max(n);
  const int w0 = i & ~63;
  if ((threadIdx.x & 63) == 0) {
    if (w0 < s.np) slice_w[w0 >> 6] = 64 * hw;
    if (err) atomicOr(&ctl->error, err);
    atomicMax(&ctl->max_nnb, hw);
  }
  if (blk == 0 && threadIdx.x == 0) ctl->nbpairs = 0;
}
__global__ __launch_bounds__(DEM_BLOCK) void k_nbr_count(SysConst s, const double4* pos,
                                                         const int* start,
                                                         const int* cell_idx, EdgeSets E,
                                                         int* slice_w, const int* gate, int force,
                                                         Ctl* ctl) {
  nbr_count_body(s, pos, start, cell_idx, E, slice_w, gate, force, ctl, blockIdx.x);
}

#define DEM_FILL_STORED 24

// The halo of tile blk of the edge set q just filled, and the staged slot of
// every edge of the tile (one block of DEM_TILE threads, one per particle).
// The neighbours outside the tile are collected in an LDS hash set, compacted
// and sorted ascending, so the halo does not depend on the order in which the
// lanes inserted them; the first E.halo_rows (at most DEM_HALO_CAP) rows are
// kept, the other neighbours get DEM_SLOT_NONE.  The slot of an
// edge is found by binary search in the sorted rows.
#define DEM_HALO_HASH 4096
#define DEM_HALO_SORT 2048
__device__ __forceinline__ void halo_build_body(const SysConst& s, const EdgeSets& E, int q,
                                                Ctl* ctl, int blk) {
  static_assert(DEM_TILE == DEM_BLOCK, "one thread per particle of a tile");
  __shared__ int h_tab[DEM_HALO_HASH];
  __shared__ int h_row[DEM_HALO_SORT];
  const int t = threadIdx.x;
  const int t0 = blk * DEM_TILE, i = t0 + t;
  for (int k = t; k < DEM_HALO_HASH; k += DEM_BLOCK) h_tab[k] = -1;
  __syncthreads();
  int deg = 0;
  long long e0 = 0;
  if (i < s.np) {
    deg = cnt_deg(E.cnt[q][i]);
    e0 = sell_edge(E.soff[q], i, 0);
  }
  const int* nbr = E.nbr[q];
  for (int m = 0; m < deg; ++m) {
    const int j = nbr[e0 + 64LL * m];
    if (j >= t0 && j < t0 + DEM_TILE) continue;
    unsigned hs = ((unsigned)j * 2654435761u) >> 20;  // 12 bits
    // at most DEM_HALO_HASH probes: a full table leaves j unstaged
    for (int probe = 0; probe < DEM_HALO_HASH; ++probe) {
      const int old = atomicCAS(&h_tab[hs], -1, j);
      if (old == -1 || old == j) break;
      hs = (hs + 1) & (DEM_HALO_HASH - 1);
    }
  }
  __syncthreads();
  constexpr int PER = DEM_HALO_HASH / DEM_BLOCK;
  int c = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) c += h_tab[t * PER + k] >= 0;
  int tot;
  int off = block_scan_excl(c, &tot);
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int v = h_tab[t * PER + k];
    if (v >= 0) {
      if (off < DEM_HALO_SORT) h_row[off] = v;
      ++off;
    }
  }
  const int n = min(tot, DEM_HALO_SORT);
  int P = 1;
  while (P < n) P <<= 1;
  for (int k = n + t; k < P; k += DEM_BLOCK) h_row[k] = 0x7fffffff;
  __syncthreads();
  for (int w = 2; w <= P; w <<= 1)
    for (int d = w >> 1; d > 0; d >>= 1) {
      for (int k = t; k < P; k += DEM_BLOCK) {
        const int l = k ^ d;
        if (l > k) {
          const bool up = (k & w) == 0;
          const int a = h_row[k], b = h_row[l];
          if ((a > b) == up) {
            h_row[k] = b;
            h_row[l] = a;
          }
        }
      }
      __syncthreads();
    }
  const int hn = min(n, E.halo_rows);
  int* halo = E.halo[q] + (size_t)blk * DEM_HALO_CAP;
  for (int k = t; k < hn; k += DEM_BLOCK) halo[k] = h_row[k];
  if (t == 0) {
    E.hcnt[q][blk] = hn;
    atomicMax(&ctl->max_halo, tot);
    if (tot > E.halo_rows) atomicAdd(&ctl->halo_over, 1);
  }
  unsigned short* slot = E.slot[q];
  for (int m = 0; m < deg; ++m) {
    const long long e = e0 + 64LL * m;
    const int j = nbr[e];
    int sl;
    if (j >= t0 && j < t0 + DEM_TILE) {
      sl = j - t0;
    } else {
      int lo = 0, hi = hn;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (h_row[mid] < j) lo = mid + 1;
        else hi = mid;
      }
      sl = (lo < hn && h_row[lo] == j) ? DEM_TILE + lo : DEM_SLOT_NONE;
    }
    slot[e] = (unsigned short)sl;
  }
}

// Pass 2 of the list build: fill the new edge set and carry the history of
// every pair that is still listed.  src maps new particle indices to the ones
// the current edge set uses (null = identity; a resort passes its
// permutation, and o2n its inverse).  Stored pairs that drop out of the list
// are counted when the reference would have kept them (cells not adjacent).
__device__ __forceinline__ void nbr_fill_body(const SysConst& s, const double4* pos,
                                              const int* start, const int* cell_idx,
                                              const EdgeSets& E, const int* src, const int* o2n,
                                              const Kin* kin_cur, double4* xref,
                                              const int* gate, int force, Ctl* ctl, int blk) {
  if (gate_closed(gate, force, ctl)) return;
  __shared__ NbrTab tab;
  stage_nbr_tab(s, tab);
  const int p = *E.epar, q = 1 - p;
  const int i = blk * DEM_BLOCK + threadIdx.x;
  const long long ecap = E.ecap;
  unsigned long long nedges = 0, dropped = 0;
  if (i < s.np) {
    const double4 P = pos[i];
    const int o = src ? src[i] : i;
    const int odeg = cnt_deg(E.cnt[p][o]);
    const long long ob = sell_edge(E.soff[p], o, 0);
    const int* onbr = E.nbr[p];
    const unsigned char* oflag = E.flag[p];
    const double2* ohist = E.hist[p];
    const int ndeg = E.cnt[q][i];
    const long long e0 = sell_edge(E.soff[q], i, 0);
    const int ci = (int)P.w;
    // Contacts first: pairs that carry history or overlap now take the first
    // slots (in scan order), the others the last ones (in reverse scan
    // order).  The force passes then find the stored history of a wavefront
    // in the first few 64-slot runs of each history plane instead of spread
    // over all of them.
    int kf = 0, kb = ndeg - 1;
    // The pairs of the old list that store history, held in registers: each
    // new neighbour is looked up there instead of in the old list in memory.
    // More than DEM_FILL_STORED of them (a dense cluster around a large
    // sphere) falls back to the search in memory.
    // (row t of the table is the t-th set bit of oldmask)
    int sj[DEM_FILL_STORED];
    int ns = 0;
    bool many = false;
#pragma unroll
    for (int t = 0; t < DEM_FILL_STORED; ++t) sj[t] = -1;
    // bit m: entry m of the old list stores history (flags requested eight at
    // a time, so the lane waits for one round trip per eight entries)
    unsigned long long oldmask = 0;
    for (int m0 = 0; m0 < odeg; m0 += 8) {
      unsigned char f[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) f[u] = (m0 + u < odeg) ? oflag[ob + 64LL * (m0 + u)] : 0;
#pragma unroll
      for (int u = 0; u < 8; ++u) oldmask |= (unsigned long long)(f[u] & 1) << (m0 + u);
    }
    // their partners, four requests at a time
    for (unsigned long long mm = oldmask; mm;) {
      int mq[4], jq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        mq[u] = mm ? __builtin_ctzll(mm) : -1;
        if (mm) mm &= mm - 1;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) jq[u] = mq[u] >= 0 ? onbr[ob + 64LL * mq[u]] : -1;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (mq[u] < 0) continue;
#pragma unroll
        for (int t = 0; t < DEM_FILL_STORED; ++t)
          if (t == ns) sj[t] = jq[u];
        if (ns == DEM_FILL_STORED) many = true;
        else ++ns;
      }
    }
    unsigned long long found = 0;  // bit m: stored entry m of the old list was carried
    // rows of the stale-pair table (none in the canonical history mode)
    const StaleTab& T = E.st;
    const int tn = (s.hist_mode == 0 && T.cnt) ? min(T.cnt[0], DEM_STALE_ROWS) : 0;
    const int id_i = tn > 0 ? kin_cur[i].id : 0;
    // one carried history in flight: edge pe of the new set, values ph*
    long long pe = -1;
    double2 ph0, ph1, ph2;
    double2* g = E.hist[q];
    auto put_pending = [&]() {
      if (pe >= 0) {
        g[pe] = ph0;
        g[ecap + pe] = ph1;
        g[2 * ecap + pe] = ph2;
      }
    };
    for_each_neighbour<2>(s, pos, i, P, start, cell_idx, tab.ok ? tab.rc2 : nullptr,
                          [&](int j, double dsq, int cj) {
      if (kf > kb) return;
      const int jo = src ? src[j] : j;
      int mo = -1;
      if (!many) {
        int tf = -1;
#pragma unroll
        for (int t = 0; t < DEM_FILL_STORED; ++t)
          if (sj[t] == jo) tf = t;
        if (tf >= 0) {
          unsigned long long mm = oldmask;
          for (int t = 0; t < tf; ++t) mm &= mm - 1;
          mo = __builtin_ctzll(mm);
        }
      } else {
        for (unsigned long long mm = oldmask; mm; mm &= mm - 1) {
          const int m = __builtin_ctzll(mm);
          if (onbr[ob + 64LL * m] == jo) {
            mo = m;
            break;
          }
        }
      }
      // a pair that comes back into the list resumes its stale history
      int tr = -1;
      if (mo < 0 && tn > 0) {
        const int r = stale_find(T.key, tn, pair_key(id_i, kin_cur[j].id));
        if (r >= 0 && T.state[r] != 2) tr = r;
      }
      const unsigned char fl = (mo >= 0 || tr >= 0) ? 1 : 0;
      if (mo >= 0) found |= 1ull << mo;
      const int pq = ci + cj * s.nc;
      const bool front = fl || (tab.ok ? tab.mindsq[pq] : s.pair[pq].mindijsq) - dsq > 0.0;
      const long long e = e0 + 64LL * (front ? kf++ : kb--);
      E.nbr[q][e] = j;
      E.flag[q][e] = fl;
      if (fl) {
        // the carried history is requested now and written when the next
        // carried pair is found (or after the scan), so the lane does not
        // wait for it here
        put_pending();
        if (mo >= 0) {
          const long long om = ob + 64LL * mo;
          ph0 = ohist[om];
          ph1 = ohist[ecap + om];
          ph2 = ohist[2 * ecap + om];
        } else {
          ph0 = T.hist[tr];
          ph1 = T.hist[DEM_STALE_ROWS + tr];
          ph2 = T.hist[2 * DEM_STALE_ROWS + tr];
          T.state[tr] = 1;  // both edges of the pair restore it
        }
        pe = e;
      }
    });
    put_pending();
    E.cnt[q][i] = ndeg | (kf << 16);
    nedges = (unsigned long long)ndeg;
    for (unsigned long long lost = oldmask & ~found; lost; lost &= lost - 1) {
      const int m = __builtin_ctzll(lost);
      const long long om = ob + 64LL * m;
      const int jn = o2n ? o2n[onbr[om]] : onbr[om];
      // A stored pair lost from the list: the reference visits it at the
      // next updating pass, and forgets it, when its cells are adjacent;
      // otherwise it keeps it, and so does the table (one row per pair,
      // written by the edge of the lower id).
      if (s.hist_mode != 0 ||
          ref_adjacent(s, kin_cur[i].rc, kin_cur[i].rxyz, kin_cur[jn].rc, kin_cur[jn].rxyz))
        continue;
      const int idi = kin_cur[i].id, idj = kin_cur[jn].id;
      if (idi > idj) continue;
      const int r = T.cnt ? T.cnt[0] + atomicAdd(&T.cnt[1], 1) : DEM_STALE_ROWS;
      if (r >= DEM_STALE_ROWS) {
        ++dropped;
        continue;
      }
      T.key[r] = pair_key(idi, idj);
      T.slot[r] = make_int2(i, jn);
      T.hist[r] = ohist[om];
      T.hist[DEM_STALE_ROWS + r] = ohist[ecap + om];
      T.hist[2 * DEM_STALE_ROWS + r] = ohist[2 * ecap + om];
      T.state[r] = 0;
    }
    xref[i] = P;
  }
  nedges = wave_sum_u64(nedges);
  dropped = wave_sum_u64(dropped);
  if ((threadIdx.x & 63) == 0) {
    if (dropped) atomicAdd(&ctl->stale_dropped, dropped);
    if (nedges) atomicAdd(&ctl->nbpairs, nedges);
  }
  if (E.halo_on) halo_build_body(s, E, q, ctl, blk);
}
__global__ __launch_bounds__(DEM_BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_nbr_fill(SysConst s, const double4* pos,
                                                        const int* start,
                                                        const int* cell_idx, EdgeSets E,
                                                        const int* src, const int* o2n,
                                                        const Kin* kin_cur, double4* xref,
                                                        const int* gate, int force, Ctl* ctl) {
  nbr_fill_body(s, pos, start, cell_idx, E, src, o2n, kin_cur, xref, gate, force, ctl,
                blockIdx.x);
}

// The halos and slots of the current edge set (when a staged launch shape
// is switched on between rebuilds).
__global__ __launch_bounds__(DEM_TILE) void k_halo_build(SysConst s, EdgeSets E, Ctl* ctl) {
  halo_build_body(s, E, *E.epar, ctl, blockIdx.x);
}

// Make the freshly built edge set current.
// The stale-pair table after a rebuild (one block of NT threads): the kept
// rows (slots through o2n when the particles were re-sorted) and the rows
// the rebuild appended, sorted by key (bitonic, in sk/si: LDS, or scratch
// memory of the system), restored and forgotten rows dropped.  Keys are
// unique, so the result does not depend on the order in which the rebuild
// appended.
#define DEM_MERGE_THREADS 1024
template <int NT>
__device__ __forceinline__ void stale_merge_body(const StaleTab& T, const int* o2n,
                                                 unsigned long long* sk, unsigned short* si) {
  constexpr int R = DEM_STALE_ROWS, PER = DEM_STALE_ROWS / NT;
  __shared__ int s_live;
  const int t = threadIdx.x;
  const int n = min(T.cnt[0], R);
  const int m = min(T.cnt[1], R - n);
  const int tot = n + m;
  if (tot == 0) return;  // the same for every thread of the block
  int P = 2;
  while (P < tot) P <<= 1;
  if (t == 0) s_live = 0;
  for (int k = t; k < P; k += NT) {
    const bool live = k < tot && (k >= n || T.state[k] == 0);
    sk[k] = live ? T.key[k] : ~0ull;
    si[k] = (unsigned short)k;
  }
  __syncthreads();
  for (int w = 2; w <= P; w <<= 1)
    for (int d = w >> 1; d > 0; d >>= 1) {
      for (int k = t; k < P; k += NT) {
        const int l = k ^ d;
        if (l > k) {
          const bool up = (k & w) == 0;
          const unsigned long long a = sk[k], b = sk[l];
          if ((a > b) == up) {
            sk[k] = b;
            sk[l] = a;
            const unsigned short x = si[k];
            si[k] = si[l];
            si[l] = x;
          }
        }
      }
      __syncthreads();
    }
  for (int k = t; k < P; k += NT)
    if (sk[k] != ~0ull && (k + 1 == P || sk[k + 1] == ~0ull)) s_live = k + 1;
  __syncthreads();
  const int live = s_live;
  unsigned long long key[PER];
  int2 slot[PER];
  double2 h0[PER], h1[PER], h2[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int k = t + r * NT;
    if (k < live) {
      const int q = si[k];
      key[r] = T.key[q];
      slot[r] = T.slot[q];
      if (q < n && o2n) slot[r] = make_int2(o2n[slot[r].x], o2n[slot[r].y]);
      h0[r] = T.hist[q];
      h1[r] = T.hist[R + q];
      h2[r] = T.hist[2 * R + q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int k = t + r * NT;
    if (k < live) {
      T.key[k] = key[r];
      T.slot[k] = slot[r];
      T.hist[k] = h0[r];
      T.hist[R + k] = h1[r];
      T.hist[2 * R + k] = h2[r];
      T.state[k] = 0;
    }
  }
  if (t == 0) {
    T.cnt[0] = live;
    T.cnt[1] = 0;
  }
}

// The flip of a rebuild and the merge of the stale-pair table it changed
// (one block of DEM_MERGE_THREADS, sorting in LDS).
__device__ __forceinline__ void nbr_flip_body(const EdgeSets& E, const int* gate, int force,
                                              Ctl* ctl, const int* o2n = nullptr) {
  __shared__ unsigned long long sk[DEM_STALE_ROWS];
  __shared__ unsigned short si[DEM_STALE_ROWS];
  if (gate_closed(gate, force, ctl)) return;
  if (threadIdx.x == 0) {
    *E.epar = 1 - *E.epar;
    ctl->rebuilds += 1;
  }
  if (E.st.cnt) stale_merge_body<DEM_MERGE_THREADS>(E.st, o2n, sk, si);
}
__global__ __launch_bounds__(DEM_MERGE_THREADS) void k_nbr_flip(EdgeSets E, const int* gate,
                                                                int force, Ctl* ctl,
                                                                const int* o2n) {
  nbr_flip_body(E, gate, force, ctl, o2n);
}

// ---------------------------------------------------------------------------
// Resort: move every per-particle array from old order to new order
// (new i <- old src[i]).  Edges are rebuilt afterwards through src.

__global__ void k_old_to_new(int np, const int* src, int* old_to_new) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) old_to_new[src[i]] = i;
}

struct PermArgs {
  const int* src;
  const double4* pos;  double4* pos2;
  const Kin* kin;      Kin* kin2;
  const double2* acc;  double2* acc2;
  const double* ring;  double* ring2;  int nring;  // 9*nprev components
  const int* id;       int* id2;
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
  a.kin2[i] = a.kin[o];
  for (int d = 0; d < 3; ++d) a.acc2[3 * (size_t)i + d] = a.acc[3 * (size_t)o + d];
  for (int c = 0; c < a.nring; ++c) a.ring2[c * N + i] = a.ring[c * N + o];
  a.id2[i] = a.id[o];
  if (a.nw > 0) {
    a.wflag2[i] = a.wflag[o];
    for (int c = 0; c < 6 * a.nw; ++c) a.whist2[c * N + i] = a.whist[c * N + o];
  }
}

// ---------------------------------------------------------------------------
// Host <-> device reordering.

// Particle records (TParticle layout) in reference order -> internal arrays
// (velocities only; id and cell state stay).
__global__ void k_aos_in(int np, const dem_particle_t* p, const int* id, double4* pos, Kin* kin) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const dem_particle_t& r = p[id[i]];
  pos[i] = make_double4(r.position[0], r.position[1], r.position[2], (double)(r.index - 1));
  kin[i].v0 = make_double2(r.velocity[0], r.velocity[1]);
  kin[i].v1 = make_double2(r.velocity[2], r.omega[0]);
  kin[i].v2 = make_double2(r.omega[1], r.omega[2]);
}

__global__ void k_aos_out(int np, dem_particle_t* p, const int* id, const double2* acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  dem_particle_t& r = p[id[i]];
  const double* a = reinterpret_cast<const double*>(acc + 3 * (size_t)i);
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

__global__ void k_cells_to_ref(int np, const Kin* kin, const int* id, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  out[id[i]] = kin[i].rc + 1;
}

// ---------------------------------------------------------------------------
// Fluid drag (src/mFluid.f90).
//
// UpdateVelocityField (:96-140) sums, for every layer z_k = -H/2 + k*Delta_Z,
// the cross-sections pi(r^2 - (z_k - z)^2) of the particles it cuts (and
// their density-weighted vertical velocities).  Each block sums its own
// particles into the layers they cut, in a fixed order; k_fluid_field adds
// the blocks in index order.  The sums are deterministic but associate
// differently from the reference's loop over particles (an ulp-level
// difference).

#define DEM_PI_REF 3.1415926535897932384626433832795

__device__ __forceinline__ double block_sum_det(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int t = threadIdx.x;
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (t == 0)
    for (int w = 0; w < DEM_BLOCK / 64; ++w) r = r + sh[w];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(DEM_BLOCK) void k_fluid_partial(SysConst s, FluidConst f,
                                                             const double4* pos, const Kin* kin,
                                                             const double* dens, double* part,
                                                             int2* range) {
  __shared__ double sh[DEM_BLOCK / 64];
  __shared__ int klo_sh[DEM_BLOCK / 64], khi_sh[DEM_BLOCK / 64];
  const int t = threadIdx.x;
  const int i = blockIdx.x * DEM_BLOCK + t;
  const bool active = i < s.np;
  double z = 0.0, r = 0.0, vz = 0.0, rho = 0.0;
  int klo = 0x7fffffff, khi = -1;
  if (active) {
    const double4 P = pos[i];
    const int c = (int)P.w;
    z = P.z;
    r = s.comp[c].radius;
    rho = dens[c];
    vz = kin[i].v1.x;
    // layers the particle may cut, one more each side for the rounding
    const double lo = (z - r - f.mh_2) / f.delta_z, hi = (z + r - f.mh_2) / f.delta_z;
    klo = lo < 0.0 ? 0 : ((int)lo > f.nl ? f.nl + 1 : (int)lo);
    khi = hi < 0.0 ? -1 : ((int)hi + 1 > f.nl ? f.nl : (int)hi + 1);
  }
  for (int o = 32; o > 0; o >>= 1) {
    klo = min(klo, __shfl_xor(klo, o, 64));
    khi = max(khi, __shfl_xor(khi, o, 64));
  }
  if ((t & 63) == 0) {
    klo_sh[t >> 6] = klo;
    khi_sh[t >> 6] = khi;
  }
  __syncthreads();
  int bl = klo_sh[0], bh = khi_sh[0];
  for (int w = 1; w < DEM_BLOCK / 64; ++w) {
    bl = min(bl, klo_sh[w]);
    bh = max(bh, khi_sh[w]);
  }
  const size_t row = (size_t)blockIdx.x * (size_t)(f.nl + 1);
  for (int k = bl; k <= bh; ++k) {
    const double zk = f.mh_2 + (double)k * f.delta_z;
    double a = 0.0, b = 0.0, c = 0.0;
    if (active) {
      const double d = zk - z;
      const double diff2 = d * d, r2 = r * r;
      if (diff2 < r2) {
        a = DEM_PI_REF * (r2 - diff2);
        const double dai = a * rho;
        b = dai * vz;
        c = dai;
      }
    }
    a = block_sum_det(a, sh);
    b = block_sum_det(b, sh);
    c = block_sum_det(c, sh);
    if (t == 0) {
      double* q = part + 3 * (row + (size_t)k);
      q[0] = a;
      q[1] = b;
      q[2] = c;
    }
  }
  if (t == 0) range[blockIdx.x] = make_int2(bl, bh);
}

// Vf(k), Ef(k) (src/mFluid.f90:128-136), one thread per layer.
__global__ void k_fluid_field(FluidConst f, int nblk, const double* part, const int2* range,
                              double* vf, double* ef) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > f.nl) return;
  double sa = 0.0, sb = 0.0, sc = 0.0;
  for (int b = 0; b < nblk; ++b) {
    const int2 rg = range[b];
    if (k < rg.x || k > rg.y) continue;
    const double* q = part + 3 * ((size_t)b * (size_t)(f.nl + 1) + (size_t)k);
    sa = sa + q[0];
    sb = sb + q[1];
    sc = sc + q[2];
  }
  if (sa == 0.0) {
    vf[k] = f.vf0;
    ef[k] = 1.0;
  } else {
    const double es = sa / f.area;
    const double e = 1.0 - es;
    ef[k] = e;
    const double vs = sb / sc;
    vf[k] = (f.vf0 - es * vs) / e;
  }
}

// DragForce (src/mFluid.f90:144-197) of every evaluated particle.
__global__ void k_fluid_drag(SysConst s, FluidConst f, const double4* pos, const Kin* kin,
                             const double* vf, const double* ef, double* drag, Ctl* ctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.np) return;
  const double4 P = pos[i];
  const Kin K = kin[i];
  double fd[3] = {0.0, 0.0, 0.0};
  const double jr = (P.z + f.h_2) / f.delta_z;
  const int j = (int)jr;
  if (!(jr >= 0.0) || j + 1 > f.nl) {
    atomicOr(&ctl->error, DEM_KERR_FLUID);
  } else {
    double vfz = vf[j];
    vfz = vfz + (jr - (double)j) * (vf[j + 1] - vfz);
    double efz = ef[j];
    efz = efz + (jr - (double)j) * (ef[j + 1] - efz);
    const d3 v = kin_v(K);
    d3 vr = neg3(v);
    vr.z = vfz + vr.z;
    const double vra = sqrt(dot3(vr, vr));
    if (vra > 0.0) {
      const double ri = s.comp[(int)P.w].radius;
      const double re = 2.0 * efz * ri * vra * f.rho_f / f.mu;
      double n = 3.6, n0 = n + 1.0, x = 0.0;
      // the fixed point converges in a few iterations; the bound keeps the
      // kernel finite for inputs where it would not
      for (int it = 0; it < 200 && fabs(n - n0) > 1.e-8; ++it) {
        n0 = n;
        x = re / pow(efz, n);
        const double sx = sqrt(x);
        const double x075 = sx * sqrt(sx);  // x**0.75 as the reference's compiler forms it
        n = (4.8 + 0.42 * x075) / (1.0 + 0.175 * x075);
      }
      double cd1 = 0.63 + 4.8 / sqrt(x);
      cd1 = cd1 * cd1;
      double cd = 0.63 + 4.8 / sqrt(re);
      cd = cd * cd;
      const double fi = (cd1 / cd) * pow(efz, 2.0 * (1.0 - n));
      const double c = cd * (DEM_PI_REF / 2.0) * ri * ri * f.rho_f * vra;
      fd[0] = c * vr.x * fi;
      fd[1] = c * vr.y * fi;
      fd[2] = c * vr.z * fi;
    }
  }
  drag[3 * (size_t)i] = fd[0];
  drag[3 * (size_t)i + 1] = fd[1];
  drag[3 * (size_t)i + 2] = fd[2];
}

// Reference cell and list rank of every particle, given by reference id: the
// cell state of the last evaluation that host-side edits keep.
__global__ void k_set_ref_cells(SysConst s, Kin* kin, const int2* cell_rank) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.np) return;
  const int2 c = cell_rank[kin[i].id];
  const int nxy = s.rn[0] * s.rn[1];
  const int z = c.x / nxy, r = c.x - z * nxy, y = r / s.rn[0], x = r - y * s.rn[0];
  kin[i].rc = c.x;
  kin[i].st = c.y;
  kin[i].rxyz = s.rc_pack ? (x | (y << 10) | (z << 20)) : 0;
}

// (cell, entry stamp) of every particle, by reference id.
__global__ void k_cell_stamps_to_ref(int np, const Kin* kin, const int* id, int2* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  out[id[i]] = make_int2(kin[i].rc, kin[i].st);
}

// Deterministic partial sums for the kinetic energy (src/MDEM.f90:199-213):
// each block sums its particles in index order, the host sums the blocks.
__global__ __launch_bounds__(DEM_BLOCK) void k_energy(SysConst s, const double4* pos,
                                                      const Kin* kin, double* part) {
  __shared__ double buf[DEM_BLOCK];
  const int i = blockIdx.x * DEM_BLOCK + threadIdx.x;
  double e = 0.0;
  if (i < s.np) {
    const CompConst c = s.comp[(int)pos[i].w];
    const Kin K = kin[i];
    const d3 v = kin_v(K), w = kin_w(K);
    e = c.mass * dot3(v, v) + c.inertia * dot3(w, w);
  }
  buf[threadIdx.x] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < DEM_BLOCK; ++k) t += buf[k];
    part[blockIdx.x] = t;
  }
}

__global__ void k_store_ring(int np, const Kin* kin, const double2* acc2, double* slot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const size_t N = (size_t)np;
  const Kin K = kin[i];
  const d3 v = kin_v(K);
  const double* acc = reinterpret_cast<const double*>(acc2);
  slot[i] = v.x;
  slot[N + i] = v.y;
  slot[2 * N + i] = v.z;
  for (int d = 0; d < 3; ++d) {
    slot[(3 + d) * N + i] = acc[6 * (size_t)i + d];
    slot[(6 + d) * N + i] = acc[6 * (size_t)i + 3 + d];
  }
}
