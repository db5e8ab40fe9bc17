// batch.hpp -- many independent systems stepped together.
//
// Every system keeps its own buffers (one dem_system each).  A batch adds a
// device table with one BatchRec per system -- its constants, the arguments
// of its two passes and of its neighbour-list rebuild, and the step
// bookkeeping (evaluation count, ABM ring order) -- uploaded once per
// dem_batch_step call and advanced on the device by k_batch_prepare at the
// start of every step, and block maps that assign each block of a launch to
// (system, block within that system).  Each phase of a step is then one
// launch for the whole batch, whatever the number and sizes of its systems,
// the launches of every step are the same, and the kernel bodies are the
// single-system ones.
#pragma once

#include "kernels.hpp"

struct RebuildArgs {
  const double4* pos;
  const Kin* kin;
  double4* xref;
  int* key;
  int* cell_count;
  int* cell_start;
  int* cursor;
  int* cell_idx;
  int* tile_sum;
  int* slice_w;
  const EdgeSets* d_es;
  EdgeSets es;
  const int* gate;
  Ctl* ctl;
  int gncell, ntiles, nslices, ntiles_np;
};

struct BatchRec {
  SysConst s;
  PassArgs p[2];  // pass 1 (forces at x_n + predictor), pass 2 (forces at the prediction + corrector)
  RebuildArgs r;
  int nblk;       // particle blocks of the system (DEM_BLOCK lanes)
  int pass_nblk;  // blocks of its force passes (the batch's launch shape)
  // step bookkeeping, advanced by k_batch_prepare
  int nprev;
  int first;      // 1 until the first step after the system joined the batch
  int step;       // steps of the current dem_batch_step call so far
  int ipc[3];     // ring slot order, 1-based (ipc(1) = most recent)
  long long eval, stamp_base;
  double* ring;
  double* psum;  // the next predictor's sums (PassArgs::psum_rd / psum_wr)
  int* rb;
  // MoveWalls of systems with WallsInMovement (nw_move > 0), at
  // Time = (next_step + step)*TimeStep before pass 2 (src/MDEM.f90:333-337)
  WallGeom* walls;
  const WallMotion* wm;
  int nw_move;
  int pad2;
  long long next_step;
  double dt, time;
};

// Arguments of the next MoveParticles step of a system, from its
// bookkeeping; the same as the host's step_args would give (evaluation
// parity, entry stamps, ring slots rotated as ipc = cshift(ipc, -1) between
// the passes, src/MDEM.f90:329-334).
__device__ __forceinline__ void prepare_step(BatchRec& r) {
  const size_t N = (size_t)r.s.np;
  const int np = r.nprev;
  int ipc[3] = {r.ipc[0], r.ipc[1], r.ipc[2]};
  long long ev = r.eval;
  PassArgs& a = r.p[0];
  int par = (int)(ev & 1);
  a.stamp_next = (int)(r.stamp_base + ev + 1);
  a.par = par;
  a.rb_clear = r.rb + par;
  a.rb_set = r.rb + (par ^ 1);
  for (int q = 0; q < np; ++q) a.slot_rd[q] = r.ring + (size_t)(ipc[q] - 1) * 9 * N;
  a.slot_wr = np > 0 ? r.ring + (size_t)(ipc[np - 1] - 1) * 9 * N : nullptr;
  a.check_pred = r.first;
  // the corrector of the previous step of this call formed the sums
  a.psum_rd = (np > 0 && r.step > 0) ? r.psum : nullptr;
  r.first = 0;
  ev += 1;
  if (np > 0) {
    const int last = ipc[np - 1];
    for (int q = np - 1; q > 0; --q) ipc[q] = ipc[q - 1];
    ipc[0] = last;
  }
  PassArgs& b = r.p[1];
  par = (int)(ev & 1);
  b.stamp_next = (int)(r.stamp_base + ev + 1);
  b.par = par;
  b.rb_clear = r.rb + par;
  b.rb_set = r.rb + (par ^ 1);
  for (int q = 0; q < np; ++q) b.slot_rd[q] = r.ring + (size_t)(ipc[q] - 1) * 9 * N;
  b.step_idx = r.step;
  b.psum_wr = np > 0 ? r.psum : nullptr;
  r.time = (double)(r.next_step + r.step) * r.dt;
  r.step += 1;
  ev += 1;
  r.r.gate = r.rb + (int)(ev & 1);
  r.eval = ev;
  for (int q = 0; q < 3; ++q) r.ipc[q] = ipc[q];
}

// The first step of a dem_batch_step call; later steps are prepared by the
// last kernel of the step before (k_nbr_flip_batch).  One thread per system.
__global__ void k_batch_prepare(BatchRec* __restrict__ rec, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) prepare_step(rec[k]);
}

// Block maps: (system, block within the system) per block of a launch.
// MAP_PASS: blocks of the force passes, of the batch's launch shape.
enum { MAP_PART = 0, MAP_CELL = 1, MAP_CTILE = 2, MAP_STILE = 3, MAP_PASS = 4, MAP_KINDS = 5 };

__device__ __forceinline__ int cell_blocks(const RebuildArgs& r) {
  return (r.gncell + 1 + DEM_BLOCK - 1) / DEM_BLOCK;
}

// map: the pass-block map of the launch shape (BS lanes per block).
template <int MODE, bool UPDATE, bool MIXED, int PI, int BS = DEM_BLOCK,
          int WAVES = DEM_PASS_WAVES, int FRONT = 1>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_pass_batch(
    const BatchRec* __restrict__ rec, const int2* __restrict__ map, int xcd_map) {
  const int2 m = map[xcd_map ? xcd_block() : (int)blockIdx.x];
  const BatchRec& r = rec[m.x];
  pass_body<MODE, UPDATE, MIXED, BS, FRONT>(r.s, r.p[PI], m.y);
}

// One block per system.
__global__ __launch_bounds__(1024) void k_pass_reduce_batch(const BatchRec* __restrict__ rec,
                                                            int pi, int update) {
  const BatchRec& r = rec[blockIdx.x];
  const PassArgs& a = r.p[pi];
  pass_reduce_body(a.ctl, a.part, min(r.pass_nblk, DEM_RED_SLOTS), a.par, update, r.s, a.kin_in,
                   r.r.es.st);
}

// One block per system.
__global__ void k_move_walls_batch(const BatchRec* __restrict__ rec) {
  const BatchRec& r = rec[blockIdx.x];
  if (r.nw_move > 0) move_walls_body(r.walls, r.wm, r.nw_move, r.time, r.r.ctl);
}

__global__ void k_bin_count_batch(const BatchRec* __restrict__ rec,
                                  const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const BatchRec& b = rec[m.x];
  bin_count_body(b.s, b.r.pos, b.r.key, b.r.cell_count, b.r.gate, 0, b.r.ctl, m.y, b.nblk);
}

// which = 0: the cell counts (map of cell tiles); 1: the slice widths (map
// of slice tiles).
__global__ __launch_bounds__(DEM_BLOCK) void k_scan_tiles_batch(const BatchRec* __restrict__ rec,
                                                                const int2* __restrict__ map,
                                                                int which) {
  const int2 m = map[blockIdx.x];
  const RebuildArgs& r = rec[m.x].r;
  if (which == 0)
    scan_tiles_body(r.cell_count, r.gncell, r.tile_sum, r.gate, 0, r.ctl, m.y);
  else
    scan_tiles_body(r.slice_w, r.nslices, r.tile_sum, r.gate, 0, r.ctl, m.y);
}

__global__ __launch_bounds__(DEM_BLOCK) void k_scan_apply_batch(const BatchRec* __restrict__ rec,
                                                                const int2* __restrict__ map,
                                                                int which) {
  const int2 m = map[blockIdx.x];
  const RebuildArgs& r = rec[m.x].r;
  if (which == 0)
    scan_apply_body(r.cell_count, r.gncell, r.tile_sum, r.cell_start, r.cursor, nullptr, 1,
                    r.gate, 0, r.ctl, m.y);
  else
    scan_apply_body(r.slice_w, r.nslices, r.tile_sum, nullptr, nullptr, r.d_es, 0, r.gate, 0,
                    r.ctl, m.y);
}

__global__ void k_bin_scatter_batch(const BatchRec* __restrict__ rec,
                                    const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const BatchRec& b = rec[m.x];
  bin_scatter_body(b.s.np, b.r.key, b.r.cursor, b.r.cell_idx, b.r.gate, 0, b.r.ctl, m.y, b.nblk);
}

__global__ void k_cell_sort_batch(const BatchRec* __restrict__ rec,
                                  const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const RebuildArgs& r = rec[m.x].r;
  cell_sort_body(r.gncell, r.cell_start, r.cell_idx, r.gate, 0, r.ctl, m.y, cell_blocks(r));
}

__global__ __launch_bounds__(DEM_BLOCK) void k_nbr_count_batch(const BatchRec* __restrict__ rec,
                                                               const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const BatchRec& b = rec[m.x];
  nbr_count_body(b.s, b.r.pos, b.r.cell_start, b.r.cell_idx, b.r.es, b.r.slice_w, b.r.gate, 0,
                 b.r.ctl, m.y);
}

__global__ __launch_bounds__(DEM_BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_nbr_fill_batch(const BatchRec* __restrict__ rec,
                                                              const int2* __restrict__ map) {
  const int2 m = map[blockIdx.x];
  const BatchRec& b = rec[m.x];
  nbr_fill_body(b.s, b.r.pos, b.r.cell_start, b.r.cell_idx, b.r.es, nullptr, nullptr, b.r.kin,
                b.r.xref, b.r.gate, 0, b.r.ctl, m.y);
}

// One block per system: the flip of the rebuild, then the arguments of the
// system's next step.
__global__ __launch_bounds__(DEM_MERGE_THREADS) void k_nbr_flip_batch(BatchRec* __restrict__ rec) {
  BatchRec& b = rec[blockIdx.x];
  nbr_flip_body(b.r.es, b.r.gate, 0, b.r.ctl);
  __syncthreads();
  if (threadIdx.x == 0) prepare_step(b);
}

// Control blocks of all systems into one array (one copy back per call).
__global__ void k_gather_ctl(const BatchRec* __restrict__ rec, int n, Ctl* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = *rec[k].r.ctl;
}

// Reset of the per-call escape bookkeeping of every system.
__global__ void k_reset_escape(const BatchRec* __restrict__ rec, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    Ctl* c = rec[k].r.ctl;
    c->stop_step = 0x7fffffff;
    c->escaped = 0;
    c->esc_pending = 0;
    c->pairs_upd_sum = 0;
    c->upd_evals = 0;
  }
}

// ---------------------------------------------------------------------------
// Persistent stepping of a batch: one cooperative launch runs a number of
// steps of every system.  The launch's blocks (all resident at once) are
// split into groups, one per system, sized by the host in proportion to the
// system's particles; a group steps its system through the phases of the
// launch-per-phase path (the same kernel bodies, in the same order, so the
// results are bitwise those of that path) and synchronises only its own
// blocks between phases.  Systems never wait for each other, and the
// rebuild phases of a step cost nothing when its system needs no rebuild.
//
// The records stay fixed during a launch: every block derives a step's
// arguments from its system's record and the step count (persist_step_vars,
// the same values prepare_step writes), so the blocks share nothing but the
// system's device state, and k_persist_advance moves the records on between
// launches.  Group barriers publish that state at agent scope (release fence,
// arrival counter, generation word, acquire fence).  Every wait is bounded:
// a block that waits past spin_limit, or sees another block give up, sets
// the abort word and leaves, and so do the others at their next barrier;
// the host then reports the failure.

struct GroupSync {
  unsigned arrive;
  unsigned gen;
};

// Scratch of the stale-pair merge of one system (DEM_STALE_ROWS rows).
struct MergeScratch {
  unsigned long long* sk;
  unsigned short* si;
  unsigned long long* key;
  int2* slot;
  double2* hist;  // three planes
};

struct PersistArgs {
  const int2* group;      // per block: (system, rank within its group)
  const int* gsize;       // per system: blocks of its group
  GroupSync* gsync;       // per system, zero at the launch
  MergeScratch* scratch;  // per system
  int* abort;
  unsigned spin_limit;
  int nsteps;
};

__device__ __forceinline__ bool group_sync(GroupSync* g, int G, unsigned& gen, int* abort,
                                           unsigned spin_limit) {
  __shared__ int s_go;
  if (G == 1) {
    // one block: its own stores are visible to it after a block barrier
    __syncthreads();
    return true;
  }
  // every wave's stores complete before the block arrives; one wave then
  // publishes them (agent-scope release) and, after the wait, makes the
  // group's stores visible to the whole block (agent-scope acquire
  // invalidates this CU's L1)
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int go = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned target = gen + 1u;
    const unsigned n =
        __hip_atomic_fetch_add(&g->arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (n == (unsigned)G) {
      __hip_atomic_store(&g->arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&g->gen, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spins = 0;
      while (__hip_atomic_load(&g->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != target) {
        if (++spins > spin_limit ||
            __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          go = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    }
    gen = target;
    if (!go) __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) go = 0;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_go = go;
  }
  __syncthreads();
  return s_go != 0;
}

// stale_merge_body for one block of DEM_BLOCK threads: sorted in the
// system's scratch, rows gathered into scratch and copied back (no row
// staging in registers).
__device__ __forceinline__ void stale_merge_scratch(const StaleTab& T, const MergeScratch& m) {
  constexpr int R = DEM_STALE_ROWS, NT = DEM_BLOCK;
  __shared__ int s_live;
  const int t = threadIdx.x;
  const int n = min(T.cnt[0], R);
  const int mm = min(T.cnt[1], R - n);
  const int tot = n + mm;
  if (tot == 0) return;
  int P = 2;
  while (P < tot) P <<= 1;
  if (t == 0) s_live = 0;
  for (int k = t; k < P; k += NT) {
    const bool live = k < tot && (k >= n || T.state[k] == 0);
    m.sk[k] = live ? T.key[k] : ~0ull;
    m.si[k] = (unsigned short)k;
  }
  __syncthreads();
  for (int w = 2; w <= P; w <<= 1)
    for (int d = w >> 1; d > 0; d >>= 1) {
      for (int k = t; k < P; k += NT) {
        const int l = k ^ d;
        if (l > k) {
          const bool up = (k & w) == 0;
          const unsigned long long a = m.sk[k], b = m.sk[l];
          if ((a > b) == up) {
            m.sk[k] = b;
            m.sk[l] = a;
            const unsigned short x = m.si[k];
            m.si[k] = m.si[l];
            m.si[l] = x;
          }
        }
      }
      __syncthreads();
    }
  for (int k = t; k < P; k += NT)
    if (m.sk[k] != ~0ull && (k + 1 == P || m.sk[k + 1] == ~0ull)) s_live = k + 1;
  __syncthreads();
  const int live = s_live;
  for (int k = t; k < live; k += NT) {
    const int q = m.si[k];
    m.key[k] = T.key[q];
    m.slot[k] = T.slot[q];
    m.hist[k] = T.hist[q];
    m.hist[R + k] = T.hist[R + q];
    m.hist[2 * R + k] = T.hist[2 * R + q];
  }
  __syncthreads();
  for (int k = t; k < live; k += NT) {
    T.key[k] = m.key[k];
    T.slot[k] = m.slot[k];
    T.hist[k] = m.hist[k];
    T.hist[R + k] = m.hist[R + k];
    T.hist[2 * R + k] = m.hist[2 * R + k];
    T.state[k] = 0;
  }
  if (t == 0) {
    T.cnt[0] = live;
    T.cnt[1] = 0;
  }
}

// The bookkeeping prepare_step advances, kept by every thread of a
// persistent block (the same values in all of them).
struct StepBook {
  long long eval;
  int ipc[3];
  int first;
  int step;
};

// The arguments of a step of a system -- what prepare_step writes into its
// record -- from the record's fixed part and the bookkeeping at the start
// of the step: pass 1 (predictor), pass 2 (corrector), the rebuild gate and
// Time after the step, and the bookkeeping of the next step.  Static
// indices only (at most three ring slots), so nothing lands in scratch.
__device__ __forceinline__ void book_ipc_shift(const BatchRec& r, StepBook& b) {
  // ipc = cshift(ipc, -1) over the first nprev slots
  if (r.nprev == 3) {
    const int x = b.ipc[2];
    b.ipc[2] = b.ipc[1];
    b.ipc[1] = b.ipc[0];
    b.ipc[0] = x;
  } else if (r.nprev == 2) {
    const int x = b.ipc[1];
    b.ipc[1] = b.ipc[0];
    b.ipc[0] = x;
  }
}
__device__ __forceinline__ StepVars book_pass1(const BatchRec& r, const StepBook& b) {
  const size_t N = (size_t)r.s.np;
  const int np = r.nprev;
  const long long ev = b.eval;
  StepVars v = step_vars(r.p[0]);
  const int par = (int)(ev & 1);
  v.stamp_next = (int)(r.stamp_base + ev + 1);
  v.par = par;
  v.rb_clear = r.rb + par;
  v.rb_set = r.rb + (par ^ 1);
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (q < np) v.slot_rd[q] = r.ring + (size_t)(b.ipc[q] - 1) * 9 * N;
  const int last = np == 3 ? b.ipc[2] : (np == 2 ? b.ipc[1] : b.ipc[0]);
  v.slot_wr = np > 0 ? r.ring + (size_t)(last - 1) * 9 * N : nullptr;
  v.check_pred = b.first;
  v.psum_rd = (np > 0 && b.step > 0) ? r.psum : nullptr;
  return v;
}
__device__ __forceinline__ StepVars book_pass2(const BatchRec& r, const StepBook& b0) {
  const size_t N = (size_t)r.s.np;
  const int np = r.nprev;
  StepBook b = b0;
  book_ipc_shift(r, b);
  const long long ev = b.eval + 1;
  StepVars v = step_vars(r.p[1]);
  const int par = (int)(ev & 1);
  v.stamp_next = (int)(r.stamp_base + ev + 1);
  v.par = par;
  v.rb_clear = r.rb + par;
  v.rb_set = r.rb + (par ^ 1);
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (q < np) v.slot_rd[q] = r.ring + (size_t)(b.ipc[q] - 1) * 9 * N;
  v.step_idx = b.step;
  v.psum_wr = np > 0 ? r.psum : nullptr;
  return v;
}
__device__ __forceinline__ const int* book_gate(const BatchRec& r, const StepBook& b) {
  return r.rb + (int)((b.eval + 2) & 1);
}
__device__ __forceinline__ double book_time(const BatchRec& r, const StepBook& b) {
  return (double)(r.next_step + b.step) * r.dt;
}
__device__ __forceinline__ void book_advance(const BatchRec& r, StepBook& b) {
  book_ipc_shift(r, b);
  b.first = 0;
  b.step += 1;
  b.eval += 2;
}

// The records after `n` steps (the launch-per-phase path's flip advances
// them step by step): one thread per system.
__global__ void k_persist_advance(BatchRec* __restrict__ rec, int ns, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ns)
    for (int s = 0; s < n; ++s) prepare_step(rec[k]);
}

// One group's phase loop over the virtual blocks [0, nv) of its system; the
// bodies keep block-local LDS state, so consecutive virtual blocks of one
// block are separated by a block barrier.
#define DEM_GROUP_FOR(v, nv) for (int v = q; v < (nv); v += G)

// Occupancy target of the persistent kernel (wavefronts per SIMD): it holds
// the step's arguments and the group's state across the pass bodies, which
// at the passes' own target would spill.
#ifndef DEM_PERSIST_WAVES
#define DEM_PERSIST_WAVES 2
#endif

template <bool MIXED, int FRONT>
__global__ __launch_bounds__(DEM_BLOCK) __attribute__((amdgpu_waves_per_eu(DEM_PERSIST_WAVES, 8)))
void k_batch_persistent(const BatchRec* __restrict__ rec, PersistArgs pa) {
  const int2 gr = pa.group[blockIdx.x];
  const int k = gr.x, q = gr.y, G = pa.gsize[k];
  const BatchRec& R = rec[k];  // fixed during the launch
  const RebuildArgs& rb = R.r;
  Ctl* ctl = rb.ctl;
  GroupSync* gs = pa.gsync + k;
  unsigned gen = 0;
  // the bookkeeping of the running step, in LDS: each phase derives its own
  // arguments from it, so none of them stays live across the pass bodies
  __shared__ StepBook sbk;
  if (threadIdx.x == 0) {
    sbk.eval = R.eval;
    for (int c = 0; c < 3; ++c) sbk.ipc[c] = R.ipc[c];
    sbk.first = R.first;
    sbk.step = R.step;
  }
  __syncthreads();
#define DEM_GROUP_SYNC() \
  if (!group_sync(gs, G, gen, pa.abort, pa.spin_limit)) return
  for (int step = 0; step < pa.nsteps; ++step) {
    // pass 1: forces at x_n + AB predictor
    DEM_GROUP_FOR(v, R.pass_nblk) {
      const StepVars v1 = book_pass1(R, sbk);
      pass_body_v<MODE_PRED, true, MIXED, DEM_BLOCK, FRONT>(R.s, R.p[0], v1, v);
      __syncthreads();
    }
    DEM_GROUP_SYNC();
    if (q == 0) {
      const StepVars v1 = book_pass1(R, sbk);
      pass_reduce_body<DEM_BLOCK>(ctl, R.p[0].part, min(R.pass_nblk, DEM_RED_SLOTS), v1.par, 1,
                                  R.s, R.p[0].kin_in, rb.es.st);
      if (R.nw_move > 0) move_walls_body(R.walls, R.wm, R.nw_move, book_time(R, sbk), ctl);
    }
    DEM_GROUP_SYNC();
    // pass 2: forces at the prediction + AM corrector
    DEM_GROUP_FOR(v, R.pass_nblk) {
      const StepVars v2 = book_pass2(R, sbk);
      pass_body_v<MODE_CORR, false, MIXED, DEM_BLOCK, FRONT>(R.s, R.p[1], v2, v);
      __syncthreads();
    }
    DEM_GROUP_SYNC();
    if (q == 0) {
      const StepVars v2 = book_pass2(R, sbk);
      pass_reduce_body<DEM_BLOCK>(ctl, R.p[1].part, min(R.pass_nblk, DEM_RED_SLOTS), v2.par, 0,
                                  R.s, R.p[1].kin_in, rb.es.st);
    }
    DEM_GROUP_SYNC();
    const int* gate = book_gate(R, sbk);
    // a step with escapes ends the system's stepping (the same in every
    // block of the group: written before the barrier, read after it)
    if (__hip_atomic_load(&ctl->escaped, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    if (__hip_atomic_load(gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      // the neighbour-list rebuild, phase by phase
      DEM_GROUP_FOR(v, R.nblk) {
        bin_count_body(R.s, rb.pos, rb.key, rb.cell_count, gate, 0, ctl, v, R.nblk);
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      DEM_GROUP_FOR(v, rb.ntiles) {
        scan_tiles_body(rb.cell_count, rb.gncell, rb.tile_sum, gate, 0, ctl, v);
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      DEM_GROUP_FOR(v, rb.ntiles) {
        scan_apply_body(rb.cell_count, rb.gncell, rb.tile_sum, rb.cell_start, rb.cursor, nullptr,
                        1, gate, 0, ctl, v);
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      DEM_GROUP_FOR(v, R.nblk) {
        bin_scatter_body(R.s.np, rb.key, rb.cursor, rb.cell_idx, gate, 0, ctl, v, R.nblk);
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      DEM_GROUP_FOR(v, cell_blocks(rb)) {
        cell_sort_body(rb.gncell, rb.cell_start, rb.cell_idx, gate, 0, ctl, v, cell_blocks(rb));
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      DEM_GROUP_FOR(v, R.nblk) {
        nbr_count_body(R.s, rb.pos, rb.cell_start, rb.cell_idx, rb.es, rb.slice_w, gate, 0, ctl, v);
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      DEM_GROUP_FOR(v, rb.ntiles_np) {
        scan_tiles_body(rb.slice_w, rb.nslices, rb.tile_sum, gate, 0, ctl, v);
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      DEM_GROUP_FOR(v, rb.ntiles_np) {
        scan_apply_body(rb.slice_w, rb.nslices, rb.tile_sum, nullptr, nullptr, rb.d_es, 0, gate, 0,
                        ctl, v);
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      DEM_GROUP_FOR(v, R.nblk) {
        nbr_fill_body(R.s, rb.pos, rb.cell_start, rb.cell_idx, rb.es, nullptr, nullptr, rb.kin,
                      rb.xref, gate, 0, ctl, v);
        __syncthreads();
      }
      DEM_GROUP_SYNC();
      if (q == 0) {
        if (threadIdx.x == 0) {
          *rb.es.epar = 1 - *rb.es.epar;
          ctl->rebuilds += 1;
        }
        if (rb.es.st.cnt) stale_merge_scratch(rb.es.st, pa.scratch[k]);
      }
      DEM_GROUP_SYNC();
    }
    __syncthreads();
    if (threadIdx.x == 0) book_advance(R, sbk);
    __syncthreads();
  }
#undef DEM_GROUP_SYNC
}
#undef DEM_GROUP_FOR
