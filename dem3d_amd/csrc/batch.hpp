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

__global__ __launch_bounds__(DEM_BLOCK) void k_nbr_fill_batch(const BatchRec* __restrict__ rec,
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
