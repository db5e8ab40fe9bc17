// dem3d.hip -- host runtime of the gfx950 DEM per-timestep path behind the
// C-ABI of include/dem3d.h.
//
// One handle owns one HIP stream and all device memory of one granular
// system.  The state stays resident; host round trips happen only in the
// upload/download calls and once at the end of dem_step (to read the control
// block: errors, escapes).
#include "../../include/dem3d.h"
#include "batch.hpp"
#include "files.hpp"
#include "observe.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

thread_local char g_err[512];

struct Timer {
  std::vector<hipEvent_t> ev;
  void ensure(size_t n) {
    while (ev.size() < n) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      ev.push_back(e);
    }
  }
  ~Timer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

}  // namespace

// Device memory of dem_observables_noalloc: one fixed array per GPU (each GPU
// holds its own copy of a __device__ variable), big enough for the partial
// sums of up to DEM_OBS_MAX_LAYERS profile layers.  Host code uses the
// device address of each GPU's copy, taken once per device in dem_create;
// the name `scratch` itself is only a host-side stand-in and is never
// passed to a kernel.
#define DEM_OBS_SCRATCH_DOUBLES                                                         \
  ((size_t)DEM_OBS_BLOCKS * DEM_OBS_NPART + (size_t)DEM_OBS_BLOCKS * DEM_OBS_MAX_LAYERS * 4 + \
   DEM_OBS_NOUT + 4 * (size_t)DEM_OBS_MAX_LAYERS)
static_assert(DEM_OBS_SCRATCH_DOUBLES * sizeof(double) == DEM_OBS_NOALLOC_BYTES,
              "DEM_OBS_NOALLOC_BYTES in dem3d.h matches the array");
static __device__ __align__(256) double scratch[DEM_OBS_SCRATCH_DOUBLES];
#define DEM_MAX_DEVICES 64
static std::mutex g_scratch_mu;               // guards g_scratch_ptr
static double* g_scratch_ptr[DEM_MAX_DEVICES];  // device address of `scratch`, per GPU
static std::mutex g_obs_noalloc_mu;           // one dem_observables_noalloc at a time

struct dem_system {
  char err[512];
  hipStream_t stream = nullptr;
  SysConst sc{};
  int order = 4, nprev = 3, n_euler = 1, hist_mode = 0;
  double dt = 0.0, skin = 0.0;
  double ab0 = 0, am0 = 0, ab[3] = {0, 0, 0}, am[3] = {0, 0, 0};
  std::vector<WallGeom> walls_h;
  std::vector<dem_wall_t> walls_in;  // as given to dem_set_walls
  std::vector<WallConst> wallc_h;
  std::vector<double> radius_h, mass_h, inertia_h;
  // device constant tables
  PairConst* d_pair = nullptr;
  CompConst* d_comp = nullptr;
  WallConst* d_wallc = nullptr;
  WallGeom* d_walls = nullptr;
  WallMotion* d_wm = nullptr;  // WM(1:nw) when WallsInMovement
  int nw_move = 0;             // nw when WallsInMovement, else 0
  double time = 0.0;           // Time (src/MConfig.f90:61)
  // particle-number changes
  int max_np = 0;              // MaxNP (src/MDEM.f90:22)
  std::vector<double> pred_w_h;  // Predicted%Index (0-based, by id) when stale
  double* d_pred_w = nullptr;
  size_t pred_w_cap = 0;
  struct Orphan {
    uint64_t key;
    double ft[3], vt[3];
  };
  std::vector<Orphan> orphans;  // stored pairs naming a particle >= NP
  int pending_cut_all = 0;      // SysConst.cut_all for the next upload (host edits)
  // fluid flow (mFluid)
  bool fluid = false;
  FluidConst fc{};
  double* d_dens = nullptr;     // Component%Density
  double* d_vf = nullptr;       // Vf(0:NL), Ef(0:NL)
  double* d_ef = nullptr;
  double* d_fpart = nullptr;    // per-block layer sums
  int2* d_frange = nullptr;     // per-block layer range
  double* d_drag = nullptr;     // drag of every particle
  size_t fpart_cap = 0, frange_cap = 0, drag_cap = 0;
  int64_t next_step = 0;       // step number of the next batched step
  // state: buffer cur is Particle, cur^1 Predicted (or Euler scratch)
  int np = 0;
  bool has_state = false;
  double4* pos[2] = {nullptr, nullptr};
  Kin* kin[2] = {nullptr, nullptr};
  int cur = 0;
  int last_eval_buf = 0;     // buffer whose kin holds the cell state of the last evaluation
  bool rc_ready = false;     // kin[cur] holds the cell state for the next evaluation
  double2* acc = nullptr;    // Particle%Acceleration/Alpha when !acc_in_ring
  double2* acc_p = nullptr;  // Predicted%Acceleration/Alpha (optional)
  double* ring = nullptr;
  double* ring2 = nullptr;
  double* psum = nullptr;  // the next predictor's ring sums (9 planes of np)
  int ipc[3] = {1, 2, 3};
  bool acc_in_ring = false;
  int* id = nullptr;
  int* id2 = nullptr;
  double4* xref = nullptr;
  EdgeSets es{};             // host copy of the two edge sets
  EdgeSets* d_es = nullptr;  // device copy (for kernels that select by *epar)
  size_t ecap = 0;           // edge slots per set (64 * slices * max_nb)
  int nslices = 0;           // 64-particle slices
  unsigned* wflag[2] = {nullptr, nullptr};
  double* whist[2] = {nullptr, nullptr};
  // binning scratch
  int* key = nullptr;
  int* cell_count = nullptr;
  int* cell_start = nullptr;
  int* cursor = nullptr;
  int* cell_idx = nullptr;
  int* tile_sum = nullptr;
  int* perm_src = nullptr;   // new -> old of the last resort
  int* old_to_new = nullptr;
  int* slice_w = nullptr;    // 64 * widest list of each slice (rebuild scratch)
  int gncell = 0, ntiles = 0, ntiles_np = 0;
  // staging
  dem_particle_t* d_aos = nullptr;
  double* d_tmp = nullptr;
  size_t tmp_bytes = 0;
  double* d_epart = nullptr;
  unsigned* part = nullptr;  // per-block pass counters
  // control
  Ctl* ctl = nullptr;
  int* rb = nullptr;  // rb[2]
  long long eval = 0;
  long long stamp_base = 0;
  long long last_update_eval = -1;
  long long steps = 0;
  int resort_interval = 500;
  long long steps_since_resort = 0;
  bool keep_pred_acc = false;
  int64_t uploaded_unlisted = 0;  // uploaded pairs with no edge in the list
  bool timing = false;
  bool mixed = false;        // DEM_MIXED force arithmetic
  int shape = 0;             // launch shape of the force passes (DEM_SHAPE_*)
  int block_map = 0;         // 0: particle ranges per XCD, 1: dispatch order
  bool in_batch = false;     // stepped by a dem_batch (its own stream must stay idle)
  dem_batch_t batch = nullptr;
  bool batch_first = false;  // next batched step is the first since joining
  Timer timer;
  double ms_pass1 = 0, ms_pass2 = 0, ms_rebuild = 0, ms_step = 0;
  int device = 0;             // the GPU of the handle's stream
  double* d_obs = nullptr;    // dem_observables partial sums
  size_t obs_cap = 0;         // doubles in d_obs
  size_t state_bytes = 0;     // device bytes of the state arrays (state_arrays)
};

static int fail(dem_handle_t h, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(h ? h->err : g_err, 512, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(h, expr)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail((h), DEM_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static hipError_t dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void**)p, n * sizeof(T));
}
template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

static inline int nblocks(long long n, int bs = DEM_BLOCK) { return (int)((n + bs - 1) / bs); }
static inline int gs_blocks(long long n) {
  long long b = (n + DEM_BLOCK - 1) / DEM_BLOCK;
  return (int)std::max(1LL, std::min(b, 4096LL));
}

static void free_state(dem_handle_t h) {
  for (int k = 0; k < 2; ++k) {
    dfree(h->pos[k]); dfree(h->kin[k]);
    dfree(h->es.cnt[k]); dfree(h->es.soff[k]);
    dfree(h->es.nbr[k]); dfree(h->es.flag[k]); dfree(h->es.hist[k]);
    dfree(h->es.slot[k]); dfree(h->es.halo[k]); dfree(h->es.hcnt[k]);
    dfree(h->wflag[k]); dfree(h->whist[k]);
  }
  dfree(h->es.epar);
  dfree(h->es.st.key); dfree(h->es.st.slot); dfree(h->es.st.hist); dfree(h->es.st.state);
  dfree(h->es.st.cnt);
  dfree(h->d_es);
  dfree(h->acc); dfree(h->acc_p); dfree(h->ring); dfree(h->ring2); dfree(h->psum); dfree(h->id); dfree(h->id2);
  dfree(h->xref); dfree(h->key); dfree(h->cell_count);
  dfree(h->cell_start); dfree(h->cursor); dfree(h->cell_idx); dfree(h->tile_sum);
  dfree(h->perm_src); dfree(h->old_to_new); dfree(h->slice_w);
  dfree(h->d_aos); dfree(h->d_tmp); dfree(h->d_epart); dfree(h->part);
  h->state_bytes = 0;
  h->np = 0;
  h->has_state = false;
  h->tmp_bytes = 0;
}

extern "C" {

const char* dem_last_error(dem_handle_t h) { return h ? h->err : g_err; }

// Device bytes of the per-handle constants and control words of dem_create.
static size_t const_bytes(int nc, int nw) {
  return (size_t)nc * nc * sizeof(PairConst) + (size_t)nc * sizeof(CompConst) +
         (size_t)std::max(1, nc * nw) * sizeof(WallConst) +
         (size_t)std::max(1, nw) * (sizeof(WallGeom) + sizeof(WallMotion)) + sizeof(Ctl) +
         2 * sizeof(int);
}

// The neighbour-search grid of a box: gn[d] cells of at least rcmax along
// each axis; returns the number of cell keys (8x8x8-cell bricks).
static long long nb_grid(const double lbox[3], double rcmax, int gn[3]) {
  long long keys = 512;
  for (int d = 0; d < 3; ++d) {
    gn[d] = std::max(1, (int)(lbox[d] / rcmax));
    keys *= (gn[d] + 7) / 8;
  }
  return keys;
}
// The neighbour-list skin of a parameter set: 0.1 of the smallest diameter
// unless given.
static double params_skin(const dem_params_t* p, double* rmax_out) {
  double rmax = 0.0, rmin = 1e300;
  for (int c = 0; c < p->nc; ++c) {
    rmax = std::max(rmax, p->radius[c]);
    rmin = std::min(rmin, p->radius[c]);
  }
  if (rmax_out) *rmax_out = rmax;
  return p->skin > 0.0 ? p->skin : 0.2 * rmin;
}

int dem_create(const dem_params_t* p, dem_handle_t* out) {
  if (!p || !out) return fail(nullptr, DEM_ERR_ARG, "null argument");
  *out = nullptr;
  if (p->nc < 1 || p->nc > DEM_MAX_COMPONENTS || p->nw < 0 || p->nw > DEM_MAX_WALLS)
    return fail(nullptr, DEM_ERR_ARG, "nc must be 1..%d and nw 0..%d", DEM_MAX_COMPONENTS,
                DEM_MAX_WALLS);
  if (p->method_order < 1 || p->method_order > 4 || p->n_euler < 1)
    return fail(nullptr, DEM_ERR_ARG, "method_order must be 1..4 and n_euler >= 1");
  if (p->max_neighbours < 0 || p->max_neighbours > 64)
    return fail(nullptr, DEM_ERR_ARG, "max_neighbours must be 0 (default 48) or 1..64");
  if (p->precision != DEM_FP64 && p->precision != DEM_MIXED)
    return fail(nullptr, DEM_ERR_ARG, "precision must be DEM_FP64 or DEM_MIXED");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1)
    return fail(nullptr, DEM_ERR_DEVICE, "no HIP device available");
  dem_handle_t h = new dem_system();
  h->err[0] = 0;
  h->es.halo_rows = DEM_HALO_CAP;
  const int nc = p->nc, nw = p->nw;
  SysConst& s = h->sc;
  s.nc = nc;
  s.nw = nw;
  s.hist_mode = h->hist_mode = p->history_mode;
  h->mixed = p->precision == DEM_MIXED;
  h->order = p->method_order;
  h->nprev = s.nprev = p->method_order - 1;
  h->n_euler = p->n_euler;
  h->dt = p->time_step;
  h->next_step = p->method_order;  // InitialStep after the start-up (src/DEM.f90:87-92)
  s.max_nb = p->max_neighbours > 0 ? p->max_neighbours : 48;
  s.max_c = 0;
  for (int c = 0; c < nc; ++c) {
    h->radius_h.push_back(p->radius[c]);
    h->mass_h.push_back(p->mass ? p->mass[c] : 0.0);
    h->inertia_h.push_back(p->inertia ? p->inertia[c] : 0.0);
  }
  double rmax = 0.0;
  h->skin = params_skin(p, &rmax);  // 0.1 of the smallest diameter unless given
  s.skin_half_sq = 0.25 * h->skin * h->skin;
  for (int d = 0; d < 3; ++d) {
    s.pbc[d] = p->pbc[d] != 0;
    s.lbox[d] = p->lbox[d];
    s.invl[d] = 1.0 / p->lbox[d];
    s.phalf[d] = 0.5 * p->lbox[d];
    s.grav[d] = p->gravity[d];
  }
  // reference grid (InitCellStructure3D, src/MCells.f90:155-227)
  int rn[3];
  if (p->ref_grid[0] > 0 && p->ref_grid[1] > 0 && p->ref_grid[2] > 0) {
    for (int d = 0; d < 3; ++d) rn[d] = p->ref_grid[d];
  } else {
    const double cutoff = 2.0 * rmax;
    const int max_ncell = p->max_ncell > 0 ? p->max_ncell : 30000;
    int n4[3];
    for (int d = 0; d < 3; ++d) n4[d] = (int)(p->lbox[d] / cutoff);
    if ((long long)n4[0] * n4[1] * n4[2] > max_ncell) {
      const double lcell =
          std::pow(p->lbox[0] * p->lbox[1] * p->lbox[2] / (double)max_ncell, 1.0 / 3.0);
      for (int d = 0; d < 3; ++d) rn[d] = (int)(p->lbox[d] / lcell);
    } else {
      for (int d = 0; d < 3; ++d) rn[d] = n4[d];
    }
  }
  for (int d = 0; d < 3; ++d) {
    if (rn[d] < 1 || (s.pbc[d] && rn[d] < 3)) {
      delete h;
      return fail(nullptr, DEM_ERR_UNSUPPORTED,
                  "reference grid has %d cells along axis %d; periodic axes need >= 3 "
                  "(with fewer the reference visits pairs twice)", rn[d], d + 1);
    }
    s.rn[d] = rn[d];
    s.rc_pack = (d == 0 ? 1 : s.rc_pack) && rn[d] <= 1024;
    s.rn_by_l[d] = (double)rn[d] / p->lbox[d];
    s.rn_by2[d] = (double)rn[d] * 0.5;
  }
  // neighbour-search grid: cells of at least 2*rmax + skin
  const long long gn = nb_grid(p->lbox, 2.0 * rmax + h->skin, s.gn);
  for (int d = 0; d < 3; ++d) {
    s.gb[d] = (s.gn[d] + 7) / 8;
    s.g_by_l[d] = (double)s.gn[d] / p->lbox[d];
    s.g_by2[d] = (double)s.gn[d] * 0.5;
  }
  if (gn > (1LL << 30)) {
    delete h;
    return fail(nullptr, DEM_ERR_ARG, "neighbour grid too large (%lld cell keys)", gn);
  }
  s.gncell = (int)gn;
  h->gncell = (int)gn;
  // ABM coefficients (InitializeAuxiliaries, src/MDEM.f90:127-154)
  const double dt = h->dt;
  double aux;
  switch (h->order) {
    case 4:
      aux = dt / 24.0;
      h->ab0 = 55.0 * aux; h->am0 = 9.0 * aux;
      h->ab[0] = -59.0 * aux; h->am[0] = 19.0 * aux;
      h->ab[1] = 37.0 * aux; h->am[1] = -5.0 * aux;
      h->ab[2] = -9.0 * aux; h->am[2] = 1.0 * aux;
      break;
    case 3:
      aux = dt / 12.0;
      h->ab0 = 23.0 * aux; h->am0 = 5.0 * aux;
      h->ab[0] = -16.0 * aux; h->am[0] = 8.0 * aux;
      h->ab[1] = 5.0 * aux; h->am[1] = -1.0 * aux;
      break;
    case 2:
      h->ab0 = 1.5 * dt; h->am0 = 0.5 * dt;
      h->ab[0] = -0.5 * dt; h->am[0] = 0.5 * dt;
      break;
    default:
      h->ab0 = dt; h->am0 = dt;
  }
  // constant tables
  std::vector<PairConst> pair(nc * nc);
  for (int t = 0; t < nc * nc; ++t) {
    PairConst& q = pair[t];
    q.kn = p->kn[t]; q.yn = p->yn[t]; q.kt = p->kt[t]; q.yt = p->yt[t]; q.musq = p->musq[t];
    q.mindij = p->min_dij[t]; q.mindijsq = p->min_dij_sq[t];
    const double rc = p->min_dij[t] + h->skin;
    q.rc2 = rc * rc;
    s.rc2_all = std::max(s.rc2_all, q.rc2);
  }
  std::vector<CompConst> comp(nc);
  for (int c = 0; c < nc; ++c) {
    CompConst& q = comp[c];
    q.radius = p->radius[c];
    q.inv_mass = p->inv_mass[c];
    q.inv_inertia = p->inv_inertia[c];
    q.eff_mass = p->eff_mass[c];
    q.mass = p->mass ? p->mass[c] : 0.0;
    q.inertia = p->inertia ? p->inertia[c] : 0.0;
    q.pad0 = q.pad1 = 0.0;
  }
  h->wallc_h.assign(std::max(1, nc * nw), WallConst{});
  for (int t = 0; t < nc * nw; ++t) {
    WallConst& q = h->wallc_h[t];
    q.knw = p->knw[t]; q.ynw = p->ynw[t]; q.ktw = p->ktw[t]; q.ytw = p->ytw[t];
    q.muwsq = p->muwsq[t];
  }
  h->walls_h.assign(std::max(1, nw), WallGeom{});
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess && (h->device < 0 || h->device >= DEM_MAX_DEVICES)) {
    delete h;
    return fail(nullptr, DEM_ERR_UNSUPPORTED, "device index beyond %d", DEM_MAX_DEVICES - 1);
  }
  if (e == hipSuccess) {
    // the device address of this GPU's copy of `scratch`, once per device
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[h->device]) {
      void* ptr = nullptr;
      e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e == hipSuccess) g_scratch_ptr[h->device] = static_cast<double*>(ptr);
    }
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = dalloc(&h->d_pair, pair.size());
  if (e == hipSuccess) e = dalloc(&h->d_comp, comp.size());
  if (e == hipSuccess) e = dalloc(&h->d_wallc, h->wallc_h.size());
  if (e == hipSuccess) e = dalloc(&h->d_walls, h->walls_h.size());
  if (e == hipSuccess) e = dalloc(&h->d_wm, h->walls_h.size());
  if (e == hipSuccess) e = dalloc(&h->ctl, 1);
  if (e == hipSuccess) e = dalloc(&h->rb, 2);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_pair, pair.data(), pair.size() * sizeof(PairConst), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_comp, comp.data(), comp.size() * sizeof(CompConst), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_wallc, h->wallc_h.data(), h->wallc_h.size() * sizeof(WallConst),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(h->ctl, 0, sizeof(Ctl));
  if (e == hipSuccess) e = hipMemset(h->rb, 0, 2 * sizeof(int));
  if (e != hipSuccess) {
    int rc = fail(nullptr, DEM_ERR_DEVICE, "device setup failed: %s", hipGetErrorString(e));
    dem_destroy(h);
    return rc;
  }
  s.pair = h->d_pair;
  s.comp = h->d_comp;
  s.wallc = h->d_wallc;
  s.walls = h->d_walls;
  *out = h;
  return DEM_OK;
}

void dem_destroy(dem_handle_t h) {
  if (h && h->batch) dem_batch_remove(h->batch, h);
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_state(h);
  dfree(h->d_pair); dfree(h->d_comp); dfree(h->d_wallc); dfree(h->d_walls); dfree(h->d_wm);
  dfree(h->d_pred_w);
  dfree(h->d_dens); dfree(h->d_vf); dfree(h->d_ef); dfree(h->d_fpart); dfree(h->d_frange);
  dfree(h->d_drag);
  dfree(h->d_obs);
  dfree(h->ctl); dfree(h->rb);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

static int sync_halo(dem_handle_t h, bool want, hipStream_t st);

int dem_set_option(dem_handle_t h, const char* name, double value) {
  if (!h || !name) return DEM_ERR_ARG;
  if (!strcmp(name, "timing")) h->timing = value != 0.0;
  else if (!strcmp(name, "resort_interval")) h->resort_interval = (int)value;
  else if (!strcmp(name, "keep_predicted_acc")) h->keep_pred_acc = value != 0.0;
  else if (!strcmp(name, "time")) h->time = value;
  else if (!strcmp(name, "next_step")) h->next_step = (int64_t)value;
  else if (!strcmp(name, "pass_shape")) {
    if (!(value >= 0.0 && value < (double)DEM_NSHAPES) || value != (double)(int)value)
      return fail(h, DEM_ERR_ARG, "pass_shape must be an integer in [0, %d)", DEM_NSHAPES);
    h->shape = (int)value;
  } else if (!strcmp(name, "block_map")) {
    if (value != 0.0 && value != 1.0) return fail(h, DEM_ERR_ARG, "block_map must be 0 or 1");
    h->block_map = (int)value;
  } else if (!strcmp(name, "halo_rows")) {
    if (!(value >= 1.0 && value <= (double)DEM_HALO_CAP) || value != (double)(int)value)
      return fail(h, DEM_ERR_ARG, "halo_rows must be an integer in [1, %d]", DEM_HALO_CAP);
    if (h->in_batch) return fail(h, DEM_ERR_STATE, "the system is stepped by a batch");
    h->es.halo_rows = (int)value;
    // halos already listed are listed again with the new bound
    if (h->es.halo_on) {
      h->es.halo_on = 0;
      const int rc = sync_halo(h, true, h->stream);
      if (rc) return rc;
    }
  } else return fail(h, DEM_ERR_ARG, "unknown option '%s'", name);
  return DEM_OK;
}

// ConfigureWalls (src/MConfig.f90:636-688).
int dem_set_walls(dem_handle_t h, int32_t nw, const dem_wall_t* w) {
  if (!h) return DEM_ERR_ARG;
  if (nw != h->sc.nw) return fail(h, DEM_ERR_ARG, "nw=%d but the system has %d walls", nw, h->sc.nw);
  if (h->in_batch) return fail(h, DEM_ERR_STATE, "the system is stepped by a batch");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int nc = h->sc.nc;
  h->walls_in.assign(w, w + nw);
  for (int j = 0; j < nw; ++j) {
    if (w[j].shape != DEM_SHAPE_RECTANGLE && w[j].shape != DEM_SHAPE_CONE)
      return fail(h, DEM_ERR_ARG, "wall %d has invalid shape %d", j + 1, w[j].shape);
    WallGeom& g = h->walls_h[j];
    g.shape = w[j].shape;
    g.l1 = w[j].l1; g.l2 = w[j].l2; g.l3 = w[j].l3;
    for (int d = 0; d < 3; ++d) {
      g.u[d] = w[j].u[d]; g.w[d] = w[j].w[d]; g.n[d] = w[j].n[d];
      g.pos[d] = w[j].position[d]; g.vel[d] = w[j].velocity[d]; g.omg[d] = w[j].omega[d];
    }
    g.rap = g.sin_t = g.cos_t = 0.0;
    for (int c = 0; c < nc; ++c) {
      WallConst& q = h->wallc_h[c + j * nc];
      q.max_alpha_u = q.min_aw_sq = q.max_aw_sq = 0.0;
    }
    if (g.shape == DEM_SHAPE_CONE) {
      const double l_2 = g.l1, l = 2.0 * l_2, ra = g.l2, rp = g.l3;
      const double aux = std::sqrt(l * l + (rp - ra) * (rp - ra));
      g.rap = aux;
      g.sin_t = l / aux;
      g.cos_t = (rp - ra) / aux;
      for (int c = 0; c < nc; ++c) {
        const double ri = h->radius_h[c];
        double mn = std::min(ra, rp) - ri;
        mn = mn > 0.0 ? mn : 0.0;
        const double mx = std::max(ra, rp) + ri;
        WallConst& q = h->wallc_h[c + j * nc];
        q.max_alpha_u = l_2 + ri;
        q.min_aw_sq = mn * mn;
        q.max_aw_sq = mx * mx;
      }
    }
  }
  HIPCHK(h, hipMemcpy(h->d_walls, h->walls_h.data(), h->walls_h.size() * sizeof(WallGeom),
                      hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->d_wallc, h->wallc_h.data(), h->wallc_h.size() * sizeof(WallConst),
                      hipMemcpyHostToDevice));
  return DEM_OK;
}

// WallsInMovement and WM(1:NW) (src/DEM.f90:716-732, scaled as :841-846).
int dem_set_wall_motion(dem_handle_t h, int32_t nw, const dem_wall_motion_t* m) {
  if (!h) return DEM_ERR_ARG;
  if (nw != 0 && nw != h->sc.nw)
    return fail(h, DEM_ERR_ARG, "nw=%d but the system has %d walls", nw, h->sc.nw);
  if (nw > 0 && !m) return fail(h, DEM_ERR_ARG, "null wall motion");
  if (h->in_batch) return fail(h, DEM_ERR_STATE, "the system is stepped by a batch");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (nw == 0) {
    h->nw_move = 0;
    return DEM_OK;
  }
  std::vector<WallMotion> wm((size_t)nw);
  for (int j = 0; j < nw; ++j) {
    for (int d = 0; d < 3; ++d) {
      // the specification refuses negative frequencies (src/DEM.f90:723-728)
      if (!(m[j].two_pi_frequency[d] >= 0.0))
        return fail(h, DEM_ERR_ARG, "wall %d: two_pi_frequency must be >= 0", j + 1);
      wm[j].cf[d] = m[j].constant_factor[d];
      wm[j].amp[d] = m[j].amplitude[d];
      wm[j].vamp[d] = m[j].velocity_amplitude[d];
      wm[j].tpf[d] = m[j].two_pi_frequency[d];
      wm[j].phase[d] = m[j].phase_angle[d];
    }
  }
  HIPCHK(h, hipMemcpy(h->d_wm, wm.data(), wm.size() * sizeof(WallMotion), hipMemcpyHostToDevice));
  h->nw_move = nw;
  return DEM_OK;
}

// FluidFlow and InitializeFluid (src/mFluid.f90:29-92).
int dem_set_fluid(dem_handle_t h, const dem_fluid_t* f) {
  if (!h) return DEM_ERR_ARG;
  if (h->in_batch) return fail(h, DEM_ERR_STATE, "the system is stepped by a batch");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (!f) {
    h->fluid = false;
    return DEM_OK;
  }
  if (f->nl < 1 || !f->density || !(f->viscosity > 0.0))
    return fail(h, DEM_ERR_ARG, "fluid needs nl >= 1, component densities and viscosity > 0");
  const SysConst& s = h->sc;
  FluidConst c{};
  c.area = s.lbox[0] * s.lbox[1];
  c.delta_z = s.lbox[2] / (double)f->nl;
  c.vf0 = f->flow_rate / c.area;
  c.h_2 = s.lbox[2] / 2.0;
  c.mh_2 = -c.h_2;
  c.rho_f = f->fluid_density;
  c.mu = f->viscosity;
  c.nl = f->nl;
  dfree(h->d_dens); dfree(h->d_vf); dfree(h->d_ef);
  h->d_dens = h->d_vf = h->d_ef = nullptr;
  HIPCHK(h, dalloc(&h->d_dens, (size_t)std::max(s.nc, 1)));
  HIPCHK(h, dalloc(&h->d_vf, (size_t)f->nl + 1));
  HIPCHK(h, dalloc(&h->d_ef, (size_t)f->nl + 1));
  HIPCHK(h, hipMemcpy(h->d_dens, f->density, (size_t)s.nc * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(h->d_vf, 0, ((size_t)f->nl + 1) * sizeof(double)));
  HIPCHK(h, hipMemset(h->d_ef, 0, ((size_t)f->nl + 1) * sizeof(double)));
  h->fc = c;
  h->fluid = true;
  return DEM_OK;
}

int dem_download_fluid_field(dem_handle_t h, int32_t nl, double* vf, double* ef) {
  if (!h || !vf || !ef) return DEM_ERR_ARG;
  if (!h->fluid || nl != h->fc.nl) return fail(h, DEM_ERR_ARG, "no fluid with %d layers", nl);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(vf, h->d_vf, ((size_t)nl + 1) * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(ef, h->d_ef, ((size_t)nl + 1) * sizeof(double), hipMemcpyDeviceToHost));
  return DEM_OK;
}

int dem_download_walls(dem_handle_t h, int32_t nw, dem_wall_t* w) {
  if (!h) return DEM_ERR_ARG;
  if (nw != h->sc.nw) return fail(h, DEM_ERR_ARG, "nw=%d but the system has %d walls", nw, h->sc.nw);
  if (nw == 0) return DEM_OK;
  if (!w) return fail(h, DEM_ERR_ARG, "null output");
  std::vector<WallGeom> g((size_t)nw);
  hipStream_t st = h->in_batch ? nullptr : h->stream;
  if (st) HIPCHK(h, hipStreamSynchronize(st));
  else HIPCHK(h, hipDeviceSynchronize());
  HIPCHK(h, hipMemcpy(g.data(), h->d_walls, g.size() * sizeof(WallGeom), hipMemcpyDeviceToHost));
  if (h->walls_in.size() != (size_t)nw) return fail(h, DEM_ERR_STATE, "walls were never set");
  for (int j = 0; j < nw; ++j) {
    w[j] = h->walls_in[j];
    for (int d = 0; d < 3; ++d) {
      w[j].position[d] = g[j].pos[d];
      w[j].velocity[d] = g[j].vel[d];
    }
  }
  return DEM_OK;
}

}  // extern "C"

static int ensure_tmp(dem_handle_t h, size_t bytes) {
  if (bytes <= h->tmp_bytes) return DEM_OK;
  dfree(h->d_tmp);
  HIPCHK(h, hipMalloc((void**)&h->d_tmp, bytes));
  h->tmp_bytes = bytes;
  return DEM_OK;
}

static int clear_contacts(dem_handle_t h);

// Sizes of a system's state for N (>= 1) particles.
static void state_dims(dem_handle_t h, size_t N) {
  h->nslices = (int)((N + 63) / 64);
  h->ecap = (size_t)h->nslices * 64 * (size_t)h->sc.max_nb;
  h->es.ecap = (long long)h->ecap;
  h->ntiles = (h->gncell + SCAN_TILE - 1) / SCAN_TILE;
  h->ntiles_np = (h->nslices + SCAN_TILE - 1) / SCAN_TILE;
}

// The device arrays of a system's state for N particles (state_dims set):
// f(&pointer, elements) for each.  alloc_state allocates them, and
// dem_footprint counts them without a device.
template <class F>
static void state_arrays(dem_handle_t h, size_t N, F&& f) {
  const SysConst& s = h->sc;
  const size_t nring = (size_t)std::max(h->nprev, 1) * 9;
  for (int k = 0; k < 2; ++k) {
    f(&h->pos[k], N);
    f(&h->kin[k], N);
    f(&h->es.cnt[k], N);
    f(&h->es.soff[k], (size_t)h->nslices + 1);
    f(&h->es.nbr[k], h->ecap);
    f(&h->es.flag[k], h->ecap);
    f(&h->es.hist[k], 3 * h->ecap);
    f(&h->es.slot[k], h->ecap);
    f(&h->es.halo[k], (size_t)nblocks(N, DEM_TILE) * DEM_HALO_CAP);
    f(&h->es.hcnt[k], (size_t)nblocks(N, DEM_TILE));
    f(&h->wflag[k], N);
    f(&h->whist[k], (size_t)std::max(s.nw, 1) * 6 * N);
  }
  f(&h->es.epar, 1);
  f(&h->es.st.key, DEM_STALE_ROWS);
  f(&h->es.st.slot, DEM_STALE_ROWS);
  f(&h->es.st.hist, 3 * DEM_STALE_ROWS);
  f(&h->es.st.state, DEM_STALE_ROWS);
  f(&h->es.st.cnt, 3);
  f(&h->d_es, 1);
  f(&h->acc, 3 * N);
  f(&h->acc_p, 3 * N);
  f(&h->ring, nring * N);
  f(&h->ring2, nring * N);
  f(&h->psum, (size_t)9 * N);
  f(&h->id, N);
  f(&h->id2, N);
  f(&h->xref, N);
  f(&h->key, N);
  f(&h->cell_count, (size_t)h->gncell + 1);
  f(&h->cell_start, (size_t)h->gncell + 1);
  f(&h->cursor, (size_t)h->gncell + 1);
  f(&h->cell_idx, N);
  f(&h->tile_sum, (size_t)std::max(h->ntiles, h->ntiles_np) + 1);
  f(&h->perm_src, N);
  f(&h->old_to_new, N);
  f(&h->slice_w, (size_t)h->nslices);
  f(&h->d_epart, (size_t)nblocks(N) + 1);
  f(&h->part, 4 * (size_t)nblocks(N, DEM_MIN_PASS_BLOCK));
}

static int alloc_state(dem_handle_t h, int np) {
  free_state(h);
  const size_t N = (size_t)std::max(np, 1);
  state_dims(h, N);
  hipError_t e = hipSuccess;
  size_t bytes = 0;
  state_arrays(h, N, [&](auto** ptr, size_t n) {
    bytes += std::max<size_t>(n, 1) * sizeof(**ptr);
    if (e == hipSuccess) e = dalloc(ptr, n);
  });
  h->state_bytes = bytes;
  // the counter slots start at zero; every reduction clears what it read
  if (e == hipSuccess)
    e = hipMemsetAsync(h->part, 0, 4 * (size_t)nblocks(N, DEM_MIN_PASS_BLOCK) * sizeof(unsigned),
                       h->stream);
  if (e != hipSuccess) {
    free_state(h);
    return fail(h, DEM_ERR_DEVICE, "allocation for %d particles failed: %s", np,
                hipGetErrorString(e));
  }
  HIPCHK(h, hipMemcpyAsync(h->d_es, &h->es, sizeof(EdgeSets), hipMemcpyHostToDevice, h->stream));
  // an empty stale-pair table (also for a state of no particles, whose
  // upload clears nothing else)
  HIPCHK(h, hipMemsetAsync(h->es.st.cnt, 0, 3 * sizeof(int), h->stream));
  HIPCHK(h, hipMemsetAsync(h->acc_p, 0, 3 * N * sizeof(double2), h->stream));
  HIPCHK(h, hipMemsetAsync(h->cell_count, 0, ((size_t)h->gncell + 1) * sizeof(int), h->stream));
  h->np = np;
  h->sc.np = np;
  // empty edge sets, current set 0 (also for a state of no particles, which
  // dem_upload_state leaves before it clears the contacts)
  return clear_contacts(h);
}

extern "C" int dem_footprint(const dem_params_t* p, int32_t np, int64_t* bytes) {
  if (!p || !bytes || np < 0) return fail(nullptr, DEM_ERR_ARG, "bad footprint arguments");
  *bytes = 0;
  if (p->nc < 1 || p->nc > DEM_MAX_COMPONENTS || p->nw < 0 || p->nw > DEM_MAX_WALLS ||
      p->method_order < 1 || p->method_order > 4 || p->max_neighbours < 0 ||
      p->max_neighbours > 64)
    return fail(nullptr, DEM_ERR_ARG, "bad parameters");
  // the state arrays alloc_state allocates, counted on a handle that holds
  // only their dimensions
  dem_system t;
  double rmax = 0.0;
  const double skin = params_skin(p, &rmax);
  t.gncell = (int)nb_grid(p->lbox, 2.0 * rmax + skin, t.sc.gn);
  t.sc.max_nb = p->max_neighbours > 0 ? p->max_neighbours : 48;
  t.sc.nw = p->nw;
  t.nprev = p->method_order - 1;
  const size_t N = (size_t)std::max(np, 1);
  state_dims(&t, N);
  size_t sum = 0;
  state_arrays(&t, N, [&](auto** ptr, size_t n) { sum += std::max<size_t>(n, 1) * sizeof(**ptr); });
  *bytes = (int64_t)(sum + const_bytes(p->nc, p->nw));
  return DEM_OK;
}

// Empty edge sets, no wall history, current set 0.
static int clear_contacts(dem_handle_t h) {
  const size_t N = (size_t)std::max(h->np, 1);
  for (int k = 0; k < 2; ++k) {
    HIPCHK(h, hipMemsetAsync(h->es.cnt[k], 0, N * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(h->es.soff[k], 0, ((size_t)h->nslices + 1) * sizeof(int),
                             h->stream));
    HIPCHK(h, hipMemsetAsync(h->wflag[k], 0, N * sizeof(unsigned), h->stream));
  }
  HIPCHK(h, hipMemsetAsync(h->es.epar, 0, sizeof(int), h->stream));
  HIPCHK(h, hipMemsetAsync(h->es.st.cnt, 0, 3 * sizeof(int), h->stream));
  return DEM_OK;
}

// ---- neighbour list ---------------------------------------------------------

// Enqueues the counting sort of particles into the neighbour grid, gated on
// *gate unless force.
static void enqueue_binning(dem_handle_t h, const double4* pos, const int* gate, int force) {
  const SysConst& s = h->sc;
  hipStream_t st = h->stream;
  const int ng = gs_blocks(h->gncell + 1);
  const int npb = gs_blocks(s.np);
  // the cell counts are zero: allocated so, and cleared by every scan of them
  hipLaunchKernelGGL(k_bin_count, dim3(npb), dim3(DEM_BLOCK), 0, st, s, pos, h->key,
                     h->cell_count, gate, force, h->ctl);
  hipLaunchKernelGGL(k_scan_tiles, dim3(h->ntiles), dim3(DEM_BLOCK), 0, st, h->cell_count,
                     h->gncell, h->tile_sum, gate, force, h->ctl);
  hipLaunchKernelGGL(k_scan_apply, dim3(h->ntiles), dim3(DEM_BLOCK), 0, st, h->cell_count,
                     h->gncell, h->tile_sum, h->cell_start, h->cursor, (const EdgeSets*)nullptr,
                     1, gate, force, h->ctl);
  hipLaunchKernelGGL(k_bin_scatter, dim3(npb), dim3(DEM_BLOCK), 0, st, s.np, h->key, h->cursor,
                     h->cell_idx, gate, force, h->ctl);
  hipLaunchKernelGGL(k_cell_sort, dim3(ng), dim3(DEM_BLOCK), 0, st, h->gncell, h->cell_start,
                     h->cell_idx, gate, force, h->ctl);
}

// Rebuild of the edge set around pos[b]: counting sort into the neighbour
// grid, count per particle, scan into the other set's offsets, fill with the
// history carried over, flip.  src/o2n translate indices when the particles
// were just re-sorted (null otherwise).
static void enqueue_rebuild(dem_handle_t h, int b, const int* gate, int force,
                            const int* src = nullptr, const int* o2n = nullptr) {
  const SysConst& s = h->sc;
  if (s.np == 0) return;
  hipStream_t st = h->stream;
  const double4* pos = h->pos[b];
  enqueue_binning(h, pos, gate, force);
  const int nb = nblocks(s.np);
  hipLaunchKernelGGL(k_nbr_count, dim3(nb), dim3(DEM_BLOCK), 0, st, s, pos, h->cell_start,
                     h->cell_idx, h->es, h->slice_w, gate, force, h->ctl);
  hipLaunchKernelGGL(k_scan_tiles, dim3(h->ntiles_np), dim3(DEM_BLOCK), 0, st, h->slice_w,
                     h->nslices, h->tile_sum, gate, force, h->ctl);
  hipLaunchKernelGGL(k_scan_apply, dim3(h->ntiles_np), dim3(DEM_BLOCK), 0, st, h->slice_w,
                     h->nslices, h->tile_sum, (int*)nullptr, (int*)nullptr, h->d_es, 0, gate,
                     force, h->ctl);
  hipLaunchKernelGGL(k_nbr_fill, dim3(nb), dim3(DEM_BLOCK), 0, st, s, pos, h->cell_start,
                     h->cell_idx, h->es, src, o2n, (const Kin*)h->kin[b], h->xref, gate, force,
                     h->ctl);
  hipLaunchKernelGGL(k_nbr_flip, dim3(1), dim3(DEM_MERGE_THREADS), 0, st, h->es, gate, force,
                     h->ctl, o2n);
}

// Re-sort the particle arrays of Particle (cur) into neighbour-grid cell
// order.  The cell state in kin[cur] moves with the particles; the edge set
// is rebuilt through the permutation so the history follows.
static int resort(dem_handle_t h) {
  SysConst& s = h->sc;
  if (s.np == 0) return DEM_OK;
  hipStream_t st = h->stream;
  enqueue_binning(h, h->pos[h->cur], h->rb, 1);
  HIPCHK(h, hipMemcpyAsync(h->perm_src, h->cell_idx, (size_t)s.np * sizeof(int),
                           hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_old_to_new, dim3(nblocks(s.np)), dim3(DEM_BLOCK), 0, st, s.np,
                     h->perm_src, h->old_to_new);
  PermArgs a;
  a.src = h->perm_src;
  a.pos = h->pos[h->cur]; a.pos2 = h->pos[h->cur ^ 1];
  a.kin = h->kin[h->cur]; a.kin2 = h->kin[h->cur ^ 1];
  a.acc = h->acc; a.acc2 = h->acc_p;
  a.ring = h->ring; a.ring2 = h->ring2; a.nring = 9 * h->nprev;
  a.id = h->id; a.id2 = h->id2;
  a.wflag = h->wflag[0]; a.wflag2 = h->wflag[1];
  a.whist = h->whist[0]; a.whist2 = h->whist[1];
  a.nw = s.nw;
  hipLaunchKernelGGL(k_permute, dim3(nblocks(s.np)), dim3(DEM_BLOCK), 0, st, s, a);
  HIPCHK(h, hipGetLastError());
  h->cur ^= 1;
  h->last_eval_buf = h->cur;
  std::swap(h->acc, h->acc_p);
  std::swap(h->ring, h->ring2);
  std::swap(h->id, h->id2);
  std::swap(h->wflag[0], h->wflag[1]);
  std::swap(h->whist[0], h->whist[1]);
  h->steps_since_resort = 0;
  enqueue_rebuild(h, h->cur, h->rb, 1, h->perm_src, h->old_to_new);
  HIPCHK(h, hipGetLastError());
  return DEM_OK;
}

static int read_ctl(dem_handle_t h, Ctl* c) {
  HIPCHK(h, hipMemcpyAsync(c, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DEM_OK;
}

static int check_errors(dem_handle_t h) {
  Ctl c;
  int rc = read_ctl(h, &c);
  if (rc) return rc;
  if (c.error) {
    std::string m;
    if (c.error & DEM_KERR_NBR_OVERFLOW)
      m += "neighbour list overflow (max_neighbours too small); ";
    if (c.error & DEM_KERR_SKIN) m += "a particle moved more than skin/2 in one step; ";
    if (c.error & DEM_KERR_ROLE) m += "overlapping pair in non-adjacent reference cells; ";
    if (c.error & DEM_KERR_CONE_AXIS) m += "particle on the axis of a cone wall; ";
    if (c.error & DEM_KERR_NAN) m += "non-finite position; ";
    if (c.error & DEM_KERR_FLUID) m += "particle outside the fluid layers; ";
    int err = c.error;
    HIPCHK(h, hipMemsetAsync(&h->ctl->error, 0, sizeof(int), h->stream));
    return fail(h, (err & DEM_KERR_NBR_OVERFLOW) ? DEM_ERR_CAPACITY : DEM_ERR_STATE, "%s",
                m.c_str());
  }
  return DEM_OK;
}

// Cell state of buffer b from its positions, against the last evaluation.
static void enqueue_refcell(dem_handle_t h, int b, int wrap = 0) {
  hipLaunchKernelGGL(k_refcell, dim3(nblocks(h->np)), dim3(DEM_BLOCK), 0, h->stream, h->sc,
                     h->pos[b], (const Kin*)h->kin[h->last_eval_buf], h->kin[b],
                     (int)(h->stamp_base + h->eval), wrap);
}

extern "C" {

// ---- upload / download ----------------------------------------------------------

int dem_upload_state(dem_handle_t h, int32_t np, const int32_t* comp, const double* x,
                     const double* v, const double* w, const double* a, const double* alpha,
                     const double* ring, const int32_t* ipc, const int32_t* cell_rank) {
  if (!h || np < 0 || (np > 0 && (!comp || !x || !v || !w)))
    return fail(h, DEM_ERR_ARG, "bad state arguments");
  if (!h->has_state || np != h->np) {
    int rc = alloc_state(h, np);
    if (rc) return rc;
  }
  const size_t N = (size_t)np;
  SysConst& s = h->sc;
  s.np = np;
  h->cur = 0;
  h->last_eval_buf = 0;
  h->acc_in_ring = false;
  h->max_np = np;
  h->pred_w_h.clear();
  h->orphans.clear();
  s.cut_all = h->pending_cut_all;
  h->pending_cut_all = 0;
  for (int k = 0; k < 3; ++k) h->ipc[k] = k + 1;
  if (ipc)
    for (int k = 0; k < h->nprev; ++k) {
      if (ipc[k] < 1 || ipc[k] > h->nprev) return fail(h, DEM_ERR_ARG, "bad ipc");
      h->ipc[k] = ipc[k];
    }
  std::vector<double4> hp(N);
  std::vector<Kin> hk(N);
  std::vector<double> ha(6 * N, 0.0);
  std::vector<int> hid(N);
  int max_rank = 0;
  for (size_t i = 0; i < N; ++i) {
    if (comp[i] < 1 || comp[i] > s.nc)
      return fail(h, DEM_ERR_ARG, "particle %zu has component %d", i + 1, comp[i]);
    hp[i] = make_double4(x[i], x[N + i], x[2 * N + i], (double)(comp[i] - 1));
    Kin& k = hk[i];
    k.v0 = make_double2(v[i], v[N + i]);
    k.v1 = make_double2(v[2 * N + i], w[i]);
    k.v2 = make_double2(w[N + i], w[2 * N + i]);
    k.id = (int)i;
    k.rc = 0;
    k.st = cell_rank ? cell_rank[i] : 0;
    k.rxyz = 0;
    for (int d = 0; d < 3; ++d) {
      if (a) ha[6 * i + d] = a[d * N + i];
      if (alpha) ha[6 * i + 3 + d] = alpha[d * N + i];
    }
    hid[i] = (int)i;
    if (cell_rank) max_rank = std::max(max_rank, cell_rank[i]);
  }
  h->has_state = true;
  h->last_update_eval = -1;
  h->rc_ready = false;
  if (np == 0) return DEM_OK;
  hipStream_t st = h->stream;
  HIPCHK(h, hipMemcpyAsync(h->pos[0], hp.data(), N * sizeof(double4), hipMemcpyHostToDevice, st));
  HIPCHK(h, hipMemcpyAsync(h->kin[0], hk.data(), N * sizeof(Kin), hipMemcpyHostToDevice, st));
  HIPCHK(h, hipMemcpyAsync((double*)h->acc, ha.data(), 6 * N * sizeof(double),
                           hipMemcpyHostToDevice, st));
  HIPCHK(h, hipMemcpyAsync(h->id, hid.data(), N * sizeof(int), hipMemcpyHostToDevice, st));
  if (h->nprev > 0) {
    if (ring)
      HIPCHK(h, hipMemcpyAsync(h->ring, ring, (size_t)h->nprev * 9 * N * sizeof(double),
                               hipMemcpyHostToDevice, st));
    else
      HIPCHK(h, hipMemsetAsync(h->ring, 0, (size_t)h->nprev * 9 * N * sizeof(double), st));
  }
  {
    int rc = clear_contacts(h);
    if (rc) return rc;
  }
  HIPCHK(h, hipMemsetAsync(h->ctl, 0, sizeof(Ctl), st));
  HIPCHK(h, hipMemsetAsync(h->rb, 0, 2 * sizeof(int), st));
  // cell state of the uploaded lists, as DistributeParticles3D
  // (src/MCells.f90:542-564) or the given order leaves them: CellIndex from
  // the positions, entry stamps = the ranks (a later transfer stamps larger)
  h->stamp_base = (long long)max_rank + 1 - h->eval;
  hipLaunchKernelGGL(k_refcell, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, st, s, h->pos[0],
                     (const Kin*)h->kin[0], h->kin[0], 0, -1);
  int rc = resort(h);
  if (rc) return rc;
  return check_errors(h);
}

}  // extern "C"

// out[c*np + ref] for c in [0, ncomp): element c of record i at
// in[i*stride_rec + c*stride_comp].
static int gather_ref(dem_handle_t h, const double* in, int ncomp, int stride_rec,
                      int stride_comp, double* host_out) {
  const size_t N = (size_t)h->np;
  int rc = ensure_tmp(h, ncomp * N * sizeof(double));
  if (rc) return rc;
  hipLaunchKernelGGL(k_to_ref_d, dim3(nblocks(h->np)), dim3(DEM_BLOCK), 0, h->stream, h->np, ncomp,
                     in, stride_rec, stride_comp, h->id, h->d_tmp);
  HIPCHK(h, hipMemcpyAsync(host_out, h->d_tmp, ncomp * N * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DEM_OK;
}

static void acc_source(dem_handle_t h, const double** p, int* rec, int* cs) {
  if (h->acc_in_ring && h->nprev > 0) {
    *p = h->ring + (size_t)(h->ipc[0] - 1) * 9 * h->np + 3 * (size_t)h->np;
    *rec = 1;
    *cs = h->np;
  } else {
    *p = (const double*)h->acc;
    *rec = 6;
    *cs = 1;
  }
}

static int download_ids(dem_handle_t h, std::vector<int>& idh) {
  idh.resize(h->np);
  if (h->np == 0) return DEM_OK;
  HIPCHK(h, hipMemcpyAsync(idh.data(), h->id, (size_t)h->np * sizeof(int), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DEM_OK;
}

// The current edge set on the host, in the device's sliced-ELL layout.
struct HostEdges {
  int par = 0;
  size_t E = 0;  // used edge slots (soff of the last slice end)
  std::vector<int> cnt, soff, nbr;
  std::vector<unsigned char> flag;
  std::vector<double2> hist;  // three planes of E entries
  size_t edge(size_t i, int k) const { return (size_t)soff[i >> 6] + 64 * (size_t)k + (i & 63); }
};

static int download_edges(dem_handle_t h, HostEdges& he, bool with_hist) {
  const size_t N = (size_t)h->np;
  const size_t ns = (size_t)h->nslices;
  hipStream_t st = h->stream;
  HIPCHK(h, hipMemcpyAsync(&he.par, h->es.epar, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  he.cnt.resize(N);
  he.soff.resize(ns + 1);
  if (N)
    HIPCHK(h, hipMemcpyAsync(he.cnt.data(), h->es.cnt[he.par], N * sizeof(int),
                             hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipMemcpyAsync(he.soff.data(), h->es.soff[he.par], (ns + 1) * sizeof(int),
                           hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  he.E = N ? (size_t)he.soff[ns] : 0;
  const size_t E = he.E;
  he.nbr.resize(E);
  he.flag.resize(E);
  if (E) {
    HIPCHK(h, hipMemcpyAsync(he.nbr.data(), h->es.nbr[he.par], E * sizeof(int),
                             hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipMemcpyAsync(he.flag.data(), h->es.flag[he.par], E, hipMemcpyDeviceToHost, st));
  }
  if (with_hist) {
    he.hist.resize(3 * E);
    for (int c = 0; c < 3 && E; ++c)
      HIPCHK(h, hipMemcpyAsync(he.hist.data() + c * E, h->es.hist[he.par] + c * h->ecap,
                               E * sizeof(double2), hipMemcpyDeviceToHost, st));
  }
  HIPCHK(h, hipStreamSynchronize(st));
  return DEM_OK;
}

extern "C" {

int dem_download_state(dem_handle_t h, int32_t np, int32_t* comp, double* x, double* v, double* w,
                       double* a, double* alpha, double* ring, int32_t* ipc) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "no state");
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np=%d but the state holds %d", np, h->np);
  const size_t N = (size_t)np;
  if (np > 0) {
    std::vector<double> tmp(4 * N);
    int rc = gather_ref(h, (const double*)h->pos[h->cur], 4, 4, 1, tmp.data());
    if (rc) return rc;
    if (x) memcpy(x, tmp.data(), 3 * N * sizeof(double));
    if (comp)
      for (size_t i = 0; i < N; ++i) comp[i] = (int)tmp[3 * N + i] + 1;
    std::vector<double> t6(6 * N);
    if (v || w) {
      rc = gather_ref(h, (const double*)h->kin[h->cur], 6, 8, 1, t6.data());
      if (rc) return rc;
      if (v) memcpy(v, t6.data(), 3 * N * sizeof(double));
      if (w) memcpy(w, t6.data() + 3 * N, 3 * N * sizeof(double));
    }
    if (a || alpha) {
      const double* src;
      int rec, cs;
      acc_source(h, &src, &rec, &cs);
      rc = gather_ref(h, src, 6, rec, cs, t6.data());
      if (rc) return rc;
      if (a) memcpy(a, t6.data(), 3 * N * sizeof(double));
      if (alpha) memcpy(alpha, t6.data() + 3 * N, 3 * N * sizeof(double));
    }
    if (ring && h->nprev > 0) {
      for (int sl = 0; sl < h->nprev; ++sl) {
        rc = gather_ref(h, h->ring + (size_t)sl * 9 * N, 9, 1, np, ring + (size_t)sl * 9 * N);
        if (rc) return rc;
      }
    }
  }
  if (ipc)
    for (int k = 0; k < h->nprev; ++k) ipc[k] = h->ipc[k];
  return DEM_OK;
}

int dem_download_predicted(dem_handle_t h, int32_t np, double* x, double* v, double* w, double* a,
                           double* alpha) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "no state");
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np mismatch");
  const size_t N = (size_t)np;
  if (np == 0) return DEM_OK;
  std::vector<double> tmp(6 * N);
  int rc = gather_ref(h, (const double*)h->pos[h->cur ^ 1], 3, 4, 1, tmp.data());
  if (rc) return rc;
  if (x) memcpy(x, tmp.data(), 3 * N * sizeof(double));
  rc = gather_ref(h, (const double*)h->kin[h->cur ^ 1], 6, 8, 1, tmp.data());
  if (rc) return rc;
  if (v) memcpy(v, tmp.data(), 3 * N * sizeof(double));
  if (w) memcpy(w, tmp.data() + 3 * N, 3 * N * sizeof(double));
  if (a || alpha) {
    rc = gather_ref(h, (const double*)h->acc_p, 6, 6, 1, tmp.data());
    if (rc) return rc;
    if (a) memcpy(a, tmp.data(), 3 * N * sizeof(double));
    if (alpha) memcpy(alpha, tmp.data() + 3 * N, 3 * N * sizeof(double));
  }
  return DEM_OK;
}

int dem_download_cells(dem_handle_t h, int32_t np, int32_t* cell_index) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "no state");
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np mismatch");
  if (np == 0) return DEM_OK;
  // CellIndex after the last evaluation
  int rc = ensure_tmp(h, (size_t)np * sizeof(int));
  if (rc) return rc;
  hipLaunchKernelGGL(k_cells_to_ref, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, h->stream, np,
                     (const Kin*)h->kin[h->last_eval_buf], h->id, (int*)h->d_tmp);
  HIPCHK(h, hipMemcpyAsync(cell_index, h->d_tmp, (size_t)np * sizeof(int), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DEM_OK;
}

// The reference's list order within each cell: head-pushed lists put the
// particle that entered last first (larger entry stamp), and particles that
// entered in the same evaluation in decreasing index order (the
// redistribution loop runs i = 1..NP, src/MCells.f90:607-643).
int dem_download_cell_rank(dem_handle_t h, int32_t np, int32_t* rank) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "no state");
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np mismatch");
  if (np == 0) return DEM_OK;
  if (!rank) return fail(h, DEM_ERR_ARG, "null output");
  int rc = ensure_tmp(h, (size_t)np * sizeof(int2));
  if (rc) return rc;
  hipLaunchKernelGGL(k_cell_stamps_to_ref, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, h->stream, np,
                     (const Kin*)h->kin[h->last_eval_buf], h->id, (int2*)h->d_tmp);
  std::vector<int2> cs((size_t)np);
  HIPCHK(h, hipMemcpyAsync(cs.data(), h->d_tmp, (size_t)np * sizeof(int2), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<int> order((size_t)np);
  for (int i = 0; i < np; ++i) order[i] = i;
  // by cell, then list position (stamp descending, index descending)
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    if (cs[a].x != cs[b].x) return cs[a].x < cs[b].x;
    if (cs[a].y != cs[b].y) return cs[a].y > cs[b].y;
    return a > b;
  });
  for (int k = 0; k < np;) {
    int e = k;
    while (e < np && cs[order[e]].x == cs[order[k]].x) ++e;
    for (int q = k; q < e; ++q) rank[order[q]] = e - 1 - q;
    k = e;
  }
  return DEM_OK;
}

// ---- files -------------------------------------------------------------------

int dem_config_write(const char* path, const dem_config_t* cfg) {
  if (!path || !cfg) return fail(nullptr, DEM_ERR_ARG, "null argument");
  demfile::Writer o;
  const std::string e = demfile::write_config(o, *cfg);
  if (!e.empty()) return fail(nullptr, DEM_ERR_ARG, "configuration: %s", e.c_str());
  if (!demfile::save(path, o)) return fail(nullptr, DEM_ERR_IO, "cannot write %s", path);
  return DEM_OK;
}

int dem_config_read(const char* path, dem_config_t* cfg) {
  if (!path || !cfg) return fail(nullptr, DEM_ERR_ARG, "null argument");
  demfile::Reader r;
  if (!demfile::load(path, r)) return fail(nullptr, DEM_ERR_IO, "cannot read %s", path);
  const std::string e = demfile::read_config(r, *cfg);
  if (!e.empty()) return fail(nullptr, DEM_ERR_IO, "%s: %s", path, e.c_str());
  return DEM_OK;
}

int dem_state_file_write(const char* path, const dem_run_t* run, const dem_config_t* cfg,
                         const dem_state_sections_t* sec) {
  if (!path || !run || !cfg || !sec) return fail(nullptr, DEM_ERR_ARG, "null argument");
  if (cfg->form != DEM_CFG_COMPLETE)
    return fail(nullptr, DEM_ERR_ARG, "a state file holds the Complete form");
  demfile::Writer o;
  std::string e = demfile::write_run_head(o, *run);
  if (e.empty()) e = demfile::write_config(o, *cfg);
  if (e.empty()) e = demfile::write_sections(o, *sec, cfg->np);
  if (!e.empty()) return fail(nullptr, DEM_ERR_ARG, "state file: %s", e.c_str());
  demfile::write_run_tail(o, *run);
  if (!demfile::save(path, o)) return fail(nullptr, DEM_ERR_IO, "cannot write %s", path);
  return DEM_OK;
}

int dem_state_file_read(const char* path, dem_run_t* run, dem_config_t* cfg,
                        dem_state_sections_t* sec) {
  if (!path || !cfg || !sec) return fail(nullptr, DEM_ERR_ARG, "null argument");
  if (sec->nprev < 0 || sec->nprev > 3) return fail(nullptr, DEM_ERR_ARG, "nprev must be 0..3");
  demfile::Reader r;
  if (!demfile::load(path, r)) return fail(nullptr, DEM_ERR_IO, "cannot read %s", path);
  dem_run_t tmp;
  memset(&tmp, 0, sizeof tmp);
  dem_run_t& rr = run ? *run : tmp;
  demfile::read_run_head(r, rr);
  if (!r.ok) return fail(nullptr, DEM_ERR_IO, "%s: file ends inside the run records", path);
  std::string e = demfile::read_config(r, *cfg);
  if (e.empty() && cfg->form != DEM_CFG_COMPLETE) e = "configuration is not in the Complete form";
  if (e.empty()) e = demfile::read_sections(r, *sec, cfg->np);
  if (e.empty()) {
    demfile::read_run_tail(r, rr);
    if (!r.ok) e = "file ends inside the run records";
  }
  if (!e.empty()) return fail(nullptr, DEM_ERR_IO, "%s: %s", path, e.c_str());
  return DEM_OK;
}

// SaveState (src/DEM.f90:947-1014) of the handle's state.
int dem_save_state_file(dem_handle_t h, const char* path, const dem_run_t* run,
                        const dem_config_t* meta) {
  if (!h || !path || !meta) return fail(h, DEM_ERR_ARG, "null argument");
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  const int np = h->np, nc = h->sc.nc, nw = h->sc.nw;
  if (nc > 0 && (!meta->density || !meta->stiffness || !meta->shear_stiffness))
    return fail(h, DEM_ERR_ARG, "meta needs density, stiffness and shear_stiffness");
  const size_t N = (size_t)np;
  std::vector<int32_t> comp(N), cells(N), rank(N), number((size_t)nc, 0), ipc(3, 0);
  std::vector<double> x(3 * N), v(3 * N), w(3 * N), ring((size_t)std::max(h->nprev, 1) * 9 * N);
  int rc = dem_download_state(h, np, comp.data(), x.data(), v.data(), w.data(), nullptr, nullptr,
                              ring.data(), ipc.data());
  if (!rc && np) rc = dem_download_cells(h, np, cells.data());
  if (!rc && np) rc = dem_download_cell_rank(h, np, rank.data());
  int64_t npairs = 0, nwall = 0;
  if (!rc)
    rc = dem_download_contacts(h, 0, &npairs, nullptr, nullptr, nullptr, 0, &nwall, nullptr,
                               nullptr, nullptr);
  std::vector<uint64_t> keys((size_t)npairs), wkeys((size_t)nwall);
  std::vector<double> ft(3 * (size_t)npairs), vt(3 * (size_t)npairs), wft(3 * (size_t)nwall),
      wvt(3 * (size_t)nwall);
  if (!rc)
    rc = dem_download_contacts(h, npairs, &npairs, keys.data(), ft.data(), vt.data(), nwall,
                               &nwall, wkeys.data(), wft.data(), wvt.data());
  std::vector<dem_wall_t> walls((size_t)std::max(nw, 1));
  if (!rc && nw) rc = dem_download_walls(h, nw, walls.data());
  dem_stats_t st;
  if (!rc) rc = dem_stats(h, &st);
  if (rc) return rc;
  // cell lists (SaveCellStructure): cells 1..Nt, each in list order
  const int ncell = h->sc.rn[0] * h->sc.rn[1] * h->sc.rn[2];
  std::vector<int32_t> len((size_t)ncell, 0), members(N);
  for (size_t i = 0; i < N; ++i) {
    if (cells[i] < 1 || cells[i] > ncell) return fail(h, DEM_ERR_STATE, "cell index out of range");
    len[cells[i] - 1] += 1;
  }
  {
    std::vector<int64_t> start((size_t)ncell + 1, 0);
    for (int c = 0; c < ncell; ++c) start[c + 1] = start[c] + len[c];
    for (size_t i = 0; i < N; ++i) {
      const int c = cells[i] - 1;
      members[(size_t)(start[c] + (len[c] - 1 - rank[i]))] = (int32_t)(i + 1);
    }
  }
  for (size_t i = 0; i < N; ++i) number[comp[i] - 1] += 1;
  dem_config_t cfg = *meta;
  cfg.form = DEM_CFG_COMPLETE;
  cfg.version = 2;
  cfg.time = h->time;
  for (int d = 0; d < 3; ++d) cfg.lbox[d] = h->sc.lbox[d];
  cfg.nc = nc;
  cfg.nw = nw;
  cfg.np = np;
  cfg.number = number.data();
  cfg.radius = h->radius_h.data();
  cfg.mass = h->mass_h.data();
  cfg.inertia = h->inertia_h.data();
  cfg.walls = walls.data();
  cfg.comp = comp.data();
  cfg.x = x.data();
  cfg.v = v.data();
  cfg.w = w.data();
  dem_state_sections_t sec;
  memset(&sec, 0, sizeof sec);
  sec.nprev = h->nprev;
  for (int k = 0; k < h->nprev; ++k) sec.ipc[k] = ipc[k];
  sec.ring = ring.data();
  sec.ncell = ncell;
  sec.cell_len = len.data();
  sec.cell_members = members.data();
  sec.npairs = npairs;
  sec.keys = keys.data();
  sec.ft = ft.data();
  sec.vt = vt.data();
  sec.nwall = nwall;
  sec.wkeys = wkeys.data();
  sec.wft = wft.data();
  sec.wvt = wvt.data();
  dem_run_t r;
  if (run) r = *run;
  else memset(&r, 0, sizeof r);
  r.ncol = (int32_t)st.ncol;
  r.next_step = (int32_t)h->next_step;
  demfile::Writer o;
  std::string e = demfile::write_run_head(o, r);
  if (e.empty()) e = demfile::write_config(o, cfg);
  if (e.empty()) e = demfile::write_sections(o, sec, np);
  if (!e.empty()) return fail(h, DEM_ERR_ARG, "state file: %s", e.c_str());
  demfile::write_run_tail(o, r);
  if (!demfile::save(path, o)) return fail(h, DEM_ERR_IO, "cannot write %s", path);
  return DEM_OK;
}

// ReadState (src/DEM.f90:1016-1110) into the handle.
int dem_load_state_file(dem_handle_t h, const char* path, dem_run_t* run, dem_config_t* meta) {
  if (!h || !path) return fail(h, DEM_ERR_ARG, "null argument");
  if (h->in_batch) return fail(h, DEM_ERR_STATE, "the system is stepped by a batch");
  demfile::Reader r;
  if (!demfile::load(path, r)) return fail(h, DEM_ERR_IO, "cannot read %s", path);
  dem_run_t rr;
  memset(&rr, 0, sizeof rr);
  demfile::read_run_head(r, rr);
  if (!r.ok) return fail(h, DEM_ERR_IO, "%s: file ends inside the run records", path);
  // sizes first, then the arrays, from the same offset
  const size_t at = r.off;
  dem_config_t c;
  memset(&c, 0, sizeof c);
  std::string e = demfile::read_config(r, c);
  if (e.empty() && c.form != DEM_CFG_COMPLETE) e = "configuration is not in the Complete form";
  if (!e.empty()) return fail(h, DEM_ERR_IO, "%s: %s", path, e.c_str());
  if (c.nc != h->sc.nc || c.nw != h->sc.nw)
    return fail(h, DEM_ERR_ARG, "%s holds %d components and %d walls, the system %d and %d", path,
                c.nc, c.nw, h->sc.nc, h->sc.nw);
  const int np = c.np, nc = c.nc, nw = c.nw;
  const size_t N = (size_t)np;
  std::vector<int32_t> number((size_t)nc), comp(N);
  std::vector<double> rad((size_t)nc), den((size_t)nc), stf((size_t)nc), shr((size_t)nc),
      ms((size_t)nc), in((size_t)nc), x(3 * N), v(3 * N), w(3 * N);
  std::vector<dem_wall_t> walls((size_t)std::max(nw, 1));
  c.number = number.data();
  c.radius = rad.data();
  c.density = den.data();
  c.stiffness = stf.data();
  c.shear_stiffness = shr.data();
  c.mass = ms.data();
  c.inertia = in.data();
  c.walls = walls.data();
  c.comp = comp.data();
  c.x = x.data();
  c.v = v.data();
  c.w = w.data();
  r.off = at;
  r.ok = true;
  e = demfile::read_config(r, c);
  const size_t at2 = r.off;
  dem_state_sections_t sec;
  memset(&sec, 0, sizeof sec);
  sec.nprev = h->nprev;
  if (e.empty()) e = demfile::read_sections(r, sec, np);
  if (!e.empty()) return fail(h, DEM_ERR_IO, "%s: %s", path, e.c_str());
  std::vector<double> ring((size_t)std::max(h->nprev, 1) * 9 * N);
  std::vector<int32_t> len((size_t)sec.ncell), members(N);
  std::vector<uint64_t> keys((size_t)sec.npairs), wkeys((size_t)sec.nwall);
  std::vector<double> ft(3 * (size_t)sec.npairs), vt(3 * (size_t)sec.npairs),
      wft(3 * (size_t)sec.nwall), wvt(3 * (size_t)sec.nwall);
  sec.ring = ring.data();
  sec.cell_len = len.data();
  sec.cell_members = members.data();
  sec.keys = keys.data();
  sec.ft = ft.data();
  sec.vt = vt.data();
  sec.wkeys = wkeys.data();
  sec.wft = wft.data();
  sec.wvt = wvt.data();
  r.off = at2;
  r.ok = true;
  e = demfile::read_sections(r, sec, np);
  if (e.empty()) {
    demfile::read_run_tail(r, rr);
    if (!r.ok) e = "file ends inside the run records";
  }
  if (!e.empty()) return fail(h, DEM_ERR_IO, "%s: %s", path, e.c_str());
  const int ncell = h->sc.rn[0] * h->sc.rn[1] * h->sc.rn[2];
  if (sec.ncell != ncell)
    return fail(h, DEM_ERR_ARG, "%s has %d cells, the system's grid %d", path, sec.ncell, ncell);
  // list position -> rank (the head of a list of n has rank n-1)
  std::vector<int32_t> rank(N, -1);
  {
    size_t m = 0;
    for (int cc = 0; cc < sec.ncell; ++cc)
      for (int k = 0; k < len[cc]; ++k, ++m) {
        const int q = members[m];
        if (q < 1 || q > np || rank[q - 1] >= 0)
          return fail(h, DEM_ERR_IO, "%s: cell lists do not hold every particle once", path);
        rank[q - 1] = len[cc] - 1 - k;
      }
  }
  int rc = DEM_OK;
  if (nw) rc = dem_set_walls(h, nw, walls.data());
  if (!rc)
    rc = dem_upload_state(h, np, comp.data(), x.data(), v.data(), w.data(), nullptr, nullptr,
                          h->nprev > 0 ? ring.data() : nullptr,
                          h->nprev > 0 ? sec.ipc : nullptr, rank.data());
  if (!rc)
    rc = dem_upload_contacts(h, sec.npairs, keys.data(), ft.data(), vt.data(), sec.nwall,
                             wkeys.data(), wft.data(), wvt.data());
  if (rc) return rc;
  h->time = c.time;
  h->next_step = rr.next_step;
  {
    const unsigned long long ncol = (unsigned long long)(uint32_t)rr.ncol;
    HIPCHK(h, hipMemcpyAsync(&h->ctl->ncol, &ncol, sizeof ncol, hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  if (run) {
    char* inter = run->interruption;
    char* ret = run->retaking;
    const int cap = run->n_interruptions;
    *run = rr;
    run->interruption = inter;
    run->retaking = ret;
    // the strings of the interruptions, when the caller gave room for them
    if ((inter || ret) && cap >= rr.n_interruptions) {
      dem_run_t again;
      memset(&again, 0, sizeof again);
      again.interruption = inter;
      again.retaking = ret;
      r.off = 0;
      r.ok = true;
      demfile::read_run_head(r, again);
    }
  }
  if (meta) {
    meta->form = c.form;
    meta->version = c.version;
    meta->length_scale = c.length_scale;
    meta->time_scale = c.time_scale;
    meta->mass_scale = c.mass_scale;
    meta->time = c.time;
    for (int d = 0; d < 3; ++d) meta->lbox[d] = c.lbox[d];
    meta->nc = nc;
    meta->nw = nw;
    meta->np = np;
    for (int k = 0; k < nc; ++k) {
      if (meta->number) meta->number[k] = number[k];
      if (meta->radius) meta->radius[k] = rad[k];
      if (meta->density) meta->density[k] = den[k];
      if (meta->stiffness) meta->stiffness[k] = stf[k];
      if (meta->shear_stiffness) meta->shear_stiffness[k] = shr[k];
      if (meta->mass) meta->mass[k] = ms[k];
      if (meta->inertia) meta->inertia[k] = in[k];
    }
    for (int j = 0; j < nw; ++j)
      if (meta->walls) meta->walls[j] = walls[j];
  }
  return DEM_OK;
}

int dem_upload_contacts(dem_handle_t h, int64_t npairs, const uint64_t* keys, const double* ft,
                        const double* vt, int64_t nwall, const uint64_t* wkeys, const double* wft,
                        const double* wvt) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  h->orphans.clear();
  const size_t N = (size_t)h->np;
  const SysConst& s = h->sc;
  std::vector<int> idh;
  int rc = download_ids(h, idh);
  if (rc) return rc;
  HostEdges he;
  rc = download_edges(h, he, false);
  if (rc) return rc;
  const size_t E = he.E;
  std::vector<unsigned char> flag(E, 0);
  std::vector<double2> hist(3 * E, make_double2(0.0, 0.0));
  std::vector<std::pair<uint64_t, int64_t>> kv((size_t)std::max<int64_t>(npairs, 0));
  for (int64_t k = 0; k < npairs; ++k) {
    const long long a = (long long)(keys[k] >> 32) - 1, b = (long long)(keys[k] & 0xffffffffu) - 1;
    if (a < 0 || b <= a || b >= (long long)N)
      return fail(h, DEM_ERR_ARG, "bad pair key at %lld", (long long)k);
    kv[k] = {keys[k], k};
  }
  std::sort(kv.begin(), kv.end());
  std::vector<unsigned char> listed(kv.size(), 0);
  for (size_t i = 0; i < N; ++i) {
    for (int m = 0; m < cnt_deg(he.cnt[i]); ++m) {
      const size_t e = he.edge(i, m);
      const int a = idh[i], b = idh[he.nbr[e]];
      const uint64_t key = pair_key(a, b);
      auto it = std::lower_bound(kv.begin(), kv.end(), std::make_pair(key, (int64_t)-1));
      if (it == kv.end() || it->first != key) continue;
      const int64_t k = it->second;
      flag[e] = 1;
      hist[e] = make_double2(ft[3 * k], ft[3 * k + 1]);
      hist[E + e] = make_double2(ft[3 * k + 2], vt[3 * k]);
      hist[2 * E + e] = make_double2(vt[3 * k + 1], vt[3 * k + 2]);
      listed[it - kv.begin()] = 1;
    }
  }
  std::vector<int> inv(N);
  for (size_t i = 0; i < N; ++i) inv[idh[i]] = (int)i;
  // A stored pair farther apart than r_i + r_j + skin has no edge: a row of
  // the stale-pair table (rows in key order), or, in the canonical history
  // mode or beyond the table's rows, dropped (counted in uploaded_unlisted).
  std::vector<unsigned long long> tkey;
  std::vector<int2> tslot;
  std::vector<double2> thist;
  int64_t unlisted = 0;
  for (size_t r = 0; r < kv.size(); ++r) {
    if (listed[r]) continue;
    if (h->hist_mode != 0 || tkey.size() == DEM_STALE_ROWS) {
      ++unlisted;
      continue;
    }
    const int64_t k = kv[r].second;
    tkey.push_back(kv[r].first);
    tslot.push_back(make_int2(inv[(kv[r].first >> 32) - 1], inv[(kv[r].first & 0xffffffffu) - 1]));
    thist.push_back(make_double2(ft[3 * k], ft[3 * k + 1]));
    thist.push_back(make_double2(ft[3 * k + 2], vt[3 * k]));
    thist.push_back(make_double2(vt[3 * k + 1], vt[3 * k + 2]));
  }
  std::vector<unsigned> wf(N, 0u);
  std::vector<double> wh((size_t)std::max(s.nw, 1) * 6 * N, 0.0);
  for (int64_t k = 0; k < nwall; ++k) {
    const long long a = (long long)(wkeys[k] >> 32) - 1, jw = (long long)(wkeys[k] & 0xffffffffu) - 1;
    if (a < 0 || a >= (long long)N || jw < 0 || jw >= s.nw)
      return fail(h, DEM_ERR_ARG, "bad wall key at %lld", (long long)k);
    const int p = inv[a];
    wf[p] |= 1u << jw;
    for (int d = 0; d < 3; ++d) {
      wh[((size_t)jw * 6 + d) * N + p] = wft[3 * k + d];
      wh[((size_t)jw * 6 + 3 + d) * N + p] = wvt[3 * k + d];
    }
  }
  hipStream_t st = h->stream;
  if (E) {
    HIPCHK(h, hipMemcpyAsync(h->es.flag[he.par], flag.data(), E, hipMemcpyHostToDevice, st));
    for (int c = 0; c < 3; ++c)
      HIPCHK(h, hipMemcpyAsync(h->es.hist[he.par] + c * h->ecap, hist.data() + c * E,
                               E * sizeof(double2), hipMemcpyHostToDevice, st));
  }
  if (N) {
    HIPCHK(h, hipMemcpyAsync(h->wflag[0], wf.data(), N * sizeof(unsigned), hipMemcpyHostToDevice,
                             st));
    if (s.nw > 0)
      HIPCHK(h, hipMemcpyAsync(h->whist[0], wh.data(), (size_t)s.nw * 6 * N * sizeof(double),
                               hipMemcpyHostToDevice, st));
  }
  const int nt = (int)tkey.size();
  if (nt) {
    std::vector<double2> planes(3 * (size_t)DEM_STALE_ROWS);
    for (int r = 0; r < nt; ++r)
      for (int c = 0; c < 3; ++c) planes[(size_t)c * DEM_STALE_ROWS + r] = thist[3 * (size_t)r + c];
    HIPCHK(h, hipMemcpyAsync(h->es.st.key, tkey.data(), nt * sizeof(unsigned long long),
                             hipMemcpyHostToDevice, st));
    HIPCHK(h, hipMemcpyAsync(h->es.st.slot, tslot.data(), nt * sizeof(int2),
                             hipMemcpyHostToDevice, st));
    HIPCHK(h, hipMemcpyAsync(h->es.st.hist, planes.data(), planes.size() * sizeof(double2),
                             hipMemcpyHostToDevice, st));
    HIPCHK(h, hipMemsetAsync(h->es.st.state, 0, nt, st));
  }
  const int tcnt[3] = {nt, 0, 0};
  HIPCHK(h, hipMemcpyAsync(h->es.st.cnt, tcnt, sizeof tcnt, hipMemcpyHostToDevice, st));
  HIPCHK(h, hipStreamSynchronize(st));
  h->uploaded_unlisted = unlisted;
  return DEM_OK;
}

int dem_download_contacts(dem_handle_t h, int64_t cap, int64_t* npairs, uint64_t* keys, double* ft,
                          double* vt, int64_t wcap, int64_t* nwall, uint64_t* wkeys, double* wft,
                          double* wvt) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "no state");
  const size_t N = (size_t)h->np;
  const SysConst& s = h->sc;
  std::vector<int> idh;
  int rc = download_ids(h, idh);
  if (rc) return rc;
  HostEdges he;
  rc = download_edges(h, he, true);
  if (rc) return rc;
  std::vector<unsigned> wf(N);
  std::vector<double> wh((size_t)std::max(s.nw, 1) * 6 * N);
  hipStream_t st = h->stream;
  if (N > 0) {
    HIPCHK(h, hipMemcpyAsync(wf.data(), h->wflag[0], N * sizeof(unsigned), hipMemcpyDeviceToHost,
                             st));
    if (s.nw > 0)
      HIPCHK(h, hipMemcpyAsync(wh.data(), h->whist[0], (size_t)s.nw * 6 * N * sizeof(double),
                               hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
  }
  struct Ent {
    uint64_t key;
    double f[3], g[3];
  };
  std::vector<Ent> pe, we;
  for (size_t i = 0; i < N; ++i) {
    const int a = idh[i];
    for (int m = 0; m < cnt_deg(he.cnt[i]); ++m) {
      const size_t e = he.edge(i, m);
      if (!(he.flag[e] & 1)) continue;
      const int b = idh[he.nbr[e]];
      if (a >= b) continue;  // each pair is stored on both of its edges
      Ent x;
      x.key = ((uint64_t)(a + 1) << 32) | (uint64_t)(b + 1);
      const double2 p0 = he.hist[e], p1 = he.hist[he.E + e], p2 = he.hist[2 * he.E + e];
      x.f[0] = p0.x; x.f[1] = p0.y; x.f[2] = p1.x;
      x.g[0] = p1.y; x.g[1] = p2.x; x.g[2] = p2.y;
      pe.push_back(x);
    }
    for (int jw = 0; jw < s.nw; ++jw) {
      if (!((wf[i] >> jw) & 1u)) continue;
      Ent x;
      x.key = ((uint64_t)(a + 1) << 32) | (uint64_t)(jw + 1);
      for (int d = 0; d < 3; ++d) {
        x.f[d] = wh[((size_t)jw * 6 + d) * N + i];
        x.g[d] = wh[((size_t)jw * 6 + 3 + d) * N + i];
      }
      we.push_back(x);
    }
  }
  for (const auto& o : h->orphans) {
    Ent x;
    x.key = o.key;
    memcpy(x.f, o.ft, sizeof x.f);
    memcpy(x.g, o.vt, sizeof x.g);
    pe.push_back(x);
  }
  // the stale pairs the table keeps
  int tcnt[3] = {0, 0, 0};
  HIPCHK(h, hipMemcpyAsync(tcnt, h->es.st.cnt, sizeof tcnt, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  if (tcnt[0] > DEM_STALE_ROWS) return fail(h, DEM_ERR_STATE, "stale-pair table corrupt");
  if (tcnt[0] > 0) {
    const size_t nt = (size_t)tcnt[0];
    std::vector<unsigned long long> tk(nt);
    std::vector<unsigned char> ts(nt);
    std::vector<double2> th(3 * (size_t)DEM_STALE_ROWS);
    HIPCHK(h, hipMemcpyAsync(tk.data(), h->es.st.key, nt * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipMemcpyAsync(ts.data(), h->es.st.state, nt, hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipMemcpyAsync(th.data(), h->es.st.hist, th.size() * sizeof(double2),
                             hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    for (size_t r = 0; r < nt; ++r) {
      if (ts[r] != 0) continue;
      const double2 p0 = th[r], p1 = th[DEM_STALE_ROWS + r], p2 = th[2 * DEM_STALE_ROWS + r];
      Ent x;
      x.key = tk[r];
      x.f[0] = p0.x; x.f[1] = p0.y; x.f[2] = p1.x;
      x.g[0] = p1.y; x.g[1] = p2.x; x.g[2] = p2.y;
      pe.push_back(x);
    }
  }
  auto by_key = [](const Ent& x, const Ent& y) { return x.key < y.key; };
  std::sort(pe.begin(), pe.end(), by_key);
  std::sort(we.begin(), we.end(), by_key);
  if (npairs) *npairs = (int64_t)pe.size();
  if (nwall) *nwall = (int64_t)we.size();
  if ((keys && (int64_t)pe.size() > cap) || (wkeys && (int64_t)we.size() > wcap))
    return fail(h, DEM_ERR_CAPACITY, "contact arrays too small");
  for (size_t k = 0; k < pe.size() && keys; ++k) {
    keys[k] = pe[k].key;
    if (ft) memcpy(ft + 3 * k, pe[k].f, 3 * sizeof(double));
    if (vt) memcpy(vt + 3 * k, pe[k].g, 3 * sizeof(double));
  }
  for (size_t k = 0; k < we.size() && wkeys; ++k) {
    wkeys[k] = we[k].key;
    if (wft) memcpy(wft + 3 * k, we[k].f, 3 * sizeof(double));
    if (wvt) memcpy(wvt + 3 * k, we[k].g, 3 * sizeof(double));
  }
  return DEM_OK;
}

}  // extern "C"

// ---- evaluation ------------------------------------------------------------------

static PassArgs base_args(dem_handle_t h) {
  PassArgs a{};
  a.xcd_map = h->block_map == 0 ? 1 : 0;
  const int p = (int)(h->eval & 1);
  a.stamp_next = (int)(h->stamp_base + h->eval + 1);
  a.par = p;
  a.xref = h->xref;
  a.E = h->es;
  a.wflag = h->wflag[0];
  a.whist = h->whist[0];
  a.rb_clear = h->rb + p;
  a.rb_set = h->rb + (p ^ 1);
  a.part = h->part;
  a.ctl = h->ctl;
  a.half_dt = 0.5 * h->dt;
  a.pred_w = h->pred_w_h.empty() ? nullptr : h->d_pred_w;
  return a;
}

// UpdateVelocityField and DragForce of the evaluated state (pos, kin) into
// h->d_drag (src/MForces.f90:170,316).
static void enqueue_fluid(dem_handle_t h, const double4* pos, const Kin* kin) {
  const int nb = nblocks(h->np);
  const size_t rows = (size_t)nb * (size_t)(h->fc.nl + 1);
  if (h->fpart_cap < 3 * rows) {
    dfree(h->d_fpart);
    h->d_fpart = nullptr;
    h->fpart_cap = dalloc(&h->d_fpart, 3 * rows) == hipSuccess ? 3 * rows : 0;
  }
  if (h->frange_cap < (size_t)nb) {
    dfree(h->d_frange);
    h->d_frange = nullptr;
    h->frange_cap = dalloc(&h->d_frange, (size_t)nb) == hipSuccess ? (size_t)nb : 0;
  }
  if (h->drag_cap < 3 * (size_t)h->np) {
    dfree(h->d_drag);
    h->d_drag = nullptr;
    h->drag_cap = dalloc(&h->d_drag, 3 * (size_t)h->np) == hipSuccess ? 3 * (size_t)h->np : 0;
  }
  if (!h->fpart_cap || !h->frange_cap || !h->drag_cap) return;
  hipStream_t st = h->stream;
  hipLaunchKernelGGL(k_fluid_partial, dim3(nb), dim3(DEM_BLOCK), 0, st, h->sc, h->fc, pos, kin,
                     (const double*)h->d_dens, h->d_fpart, h->d_frange);
  hipLaunchKernelGGL(k_fluid_field, dim3(nblocks(h->fc.nl + 1, 64)), dim3(64), 0, st, h->fc, nb,
                     (const double*)h->d_fpart, (const int2*)h->d_frange, h->d_vf, h->d_ef);
  hipLaunchKernelGGL(k_fluid_drag, dim3(nb), dim3(DEM_BLOCK), 0, st, h->sc, h->fc, pos, kin,
                     (const double*)h->d_vf, (const double*)h->d_ef, h->d_drag, h->ctl);
}

template <int S = 0>
static void shape_table(int* block, int* waves, int* front) {
  if constexpr (S < DEM_NSHAPES) {
    block[S] = PassShape<S>::block;
    waves[S] = PassShape<S>::waves;
    front[S] = PassShape<S>::front;
    shape_table<S + 1>(block, waves, front);
  }
}
static int shape_block(int shape) {
  int b[DEM_NSHAPES], w[DEM_NSHAPES], d[DEM_NSHAPES];
  shape_table(b, w, d);
  return b[shape];
}
static bool shape_staged(int shape) {
  int b[DEM_NSHAPES], w[DEM_NSHAPES], d[DEM_NSHAPES];
  shape_table(b, w, d);
  return (d[shape] & DEM_PV_STAGED) != 0;
}

// Tile halos are listed by the neighbour-list rebuilds only while a staged
// launch shape is in use: switching one on lists the halos of the current
// lists (k_halo_build, on stream st), switching it off stops the listing.
static int sync_halo(dem_handle_t h, bool want, hipStream_t st) {
  if (want == (h->es.halo_on != 0)) return DEM_OK;
  h->es.halo_on = want ? 1 : 0;
  HIPCHK(h, hipMemcpyAsync(h->d_es, &h->es, sizeof(EdgeSets), hipMemcpyHostToDevice, st));
  if (want && h->np > 0)
    hipLaunchKernelGGL(k_halo_build, dim3(nblocks(h->np, DEM_TILE)), dim3(DEM_TILE), 0, st,
                       h->sc, h->es, h->ctl);
  HIPCHK(h, hipGetLastError());
  return DEM_OK;
}

// The pass kernel in launch shape `shape` (nb blocks).  Only the steps'
// passes (predictor, corrector) come in every shape; the others run in the
// default one.
template <int MODE, bool UPDATE, bool MIXED, int S = 0>
static void launch_shape(int shape, int nb, hipStream_t st, const SysConst& s, const PassArgs& a) {
  if constexpr ((MODE == MODE_PRED || MODE == MODE_CORR) && S < DEM_NSHAPES) {
    using Sh = PassShape<S>;
    if (shape == S) {
      hipLaunchKernelGGL((k_pass<MODE, UPDATE, MIXED, Sh::block, Sh::waves, Sh::front>), dim3(nb),
                         dim3(Sh::block), 0, st, s, a);
      return;
    }
    launch_shape<MODE, UPDATE, MIXED, S + 1>(shape, nb, st, s, a);
  } else {
    hipLaunchKernelGGL((k_pass<MODE, UPDATE, MIXED>), dim3(nb), dim3(DEM_BLOCK), 0, st, s, a);
  }
}

// One force evaluation: the pass kernel, then the sum of its counters.
template <int MODE, bool UPDATE>
static void launch_pass(dem_handle_t h, const PassArgs& a0) {
  PassArgs a = a0;
  if (h->fluid) {
    enqueue_fluid(h, a.pos_in, a.kin_in);
    a.drag = h->drag_cap ? h->d_drag : nullptr;
  }
  const int nb = (MODE == MODE_PRED || MODE == MODE_CORR) ? nblocks(h->np, shape_block(h->shape))
                                                          : nblocks(h->np);
  if (h->mixed)
    launch_shape<MODE, UPDATE, true>(h->shape, nb, h->stream, h->sc, a);
  else
    launch_shape<MODE, UPDATE, false>(h->shape, nb, h->stream, h->sc, a);
  hipLaunchKernelGGL(k_pass_reduce, dim3(1), dim3(1024), 0, h->stream, h->ctl, a.part, nb, a.par,
                     UPDATE ? 1 : 0, h->sc, a.kin_in, h->es.st);
}

static void after_eval(dem_handle_t h, bool update, int b) {
  if (update) h->last_update_eval = h->eval;
  h->last_eval_buf = b;
  h->eval += 1;
}

extern "C" {

int dem_compute_accelerations_aos(dem_handle_t h, int32_t np, dem_particle_t* p, double time_step,
                                  int32_t update) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (np != h->np) return fail(h, DEM_ERR_STATE, "state holds %d particles, call passes %d", h->np, np);
  if (np == 0) return DEM_OK;
  const size_t N = (size_t)np;
  hipStream_t st = h->stream;
  if (!h->d_aos) HIPCHK(h, hipMalloc((void**)&h->d_aos, N * sizeof(dem_particle_t)));
  HIPCHK(h, hipMemcpyAsync(h->d_aos, p, N * sizeof(dem_particle_t), hipMemcpyHostToDevice, st));
  const int b = h->cur;
  hipLaunchKernelGGL(k_aos_in, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, st, np, h->d_aos, h->id,
                     h->pos[b], h->kin[b]);
  enqueue_refcell(h, b);
  const int par = (int)(h->eval & 1);
  hipLaunchKernelGGL(k_check_disp, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, st, h->sc, h->pos[b],
                     h->xref, h->rb + par);
  enqueue_rebuild(h, b, h->rb + par, 0);
  PassArgs a = base_args(h);
  a.pos_in = h->pos[b];
  a.kin_in = h->kin[b];
  a.acc_out = h->acc;
  a.half_dt = 0.5 * time_step;
  if (update)
    launch_pass<MODE_ACC, true>(h, a);
  else
    launch_pass<MODE_ACC, false>(h, a);
  HIPCHK(h, hipGetLastError());
  after_eval(h, update != 0, b);
  h->rc_ready = false;
  h->acc_in_ring = false;
  hipLaunchKernelGGL(k_aos_out, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, st, np, h->d_aos, h->id,
                     (const double2*)h->acc);
  HIPCHK(h, hipMemcpyAsync(p, h->d_aos, N * sizeof(dem_particle_t), hipMemcpyDeviceToHost, st));
  return check_errors(h);
}

int dem_compute_accelerations(dem_handle_t h, int32_t which, int32_t update) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (h->np == 0) return DEM_OK;
  const int np = h->np;
  hipStream_t st = h->stream;
  const int b = which ? (h->cur ^ 1) : h->cur;
  enqueue_refcell(h, b);
  const int par = (int)(h->eval & 1);
  hipLaunchKernelGGL(k_check_disp, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, st, h->sc, h->pos[b],
                     h->xref, h->rb + par);
  enqueue_rebuild(h, b, h->rb + par, 0);
  PassArgs a = base_args(h);
  a.pos_in = h->pos[b];
  a.kin_in = h->kin[b];
  a.acc_out = which ? h->acc_p : h->acc;
  if (update)
    launch_pass<MODE_ACC, true>(h, a);
  else
    launch_pass<MODE_ACC, false>(h, a);
  HIPCHK(h, hipGetLastError());
  after_eval(h, update != 0, b);
  h->rc_ready = false;
  if (!which) h->acc_in_ring = false;
  return check_errors(h);
}

// MoveWalls(time) on the handle's stream when WallsInMovement.
static void enqueue_move_walls(dem_handle_t h, double time) {
  if (h->nw_move > 0)
    hipLaunchKernelGGL(k_move_walls, dim3(1), dim3(64), 0, h->stream, h->d_walls,
                       (const WallMotion*)h->d_wm, h->nw_move, time, (const Ctl*)h->ctl);
}

int dem_euler_startup(dem_handle_t h) {
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (h->nprev == 0) return fail(h, DEM_ERR_STATE, "Euler start-up needs MethodOrder > 1");
  if (h->acc_in_ring)
    return fail(h, DEM_ERR_STATE, "Euler start-up after MoveParticles is not supported");
  if (h->np == 0) {
    // no particles yet (a box to be poured): the ring order, Time and the
    // walls advance all the same
    const int last = h->ipc[h->nprev - 1];
    for (int k = h->nprev - 1; k > 0; --k) h->ipc[k] = h->ipc[k - 1];
    h->ipc[0] = last;
    const double hs = h->dt / (double)h->n_euler;
    for (int sub = 0; sub < h->n_euler; ++sub) {
      h->time = h->time + hs;
      enqueue_move_walls(h, h->time);
    }
    HIPCHK(h, hipGetLastError());
    return DEM_OK;
  }
  const int np = h->np;
  hipStream_t st = h->stream;
  // ipc = cshift(ipc, -1); Previous(ipc(1)) = (v, a, alpha)  (src/MDEM.f90:266-271)
  {
    const int last = h->ipc[h->nprev - 1];
    for (int k = h->nprev - 1; k > 0; --k) h->ipc[k] = h->ipc[k - 1];
    h->ipc[0] = last;
  }
  hipLaunchKernelGGL(k_store_ring, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, st, np,
                     (const Kin*)h->kin[h->cur], (const double2*)h->acc,
                     h->ring + (size_t)(h->ipc[0] - 1) * 9 * np);
  if (!h->rc_ready) enqueue_refcell(h, h->cur);
  const double hs = h->dt / (double)h->n_euler;
  for (int sub = 0; sub < h->n_euler; ++sub) {
    PassArgs a = base_args(h);
    a.pos_in = h->pos[h->cur];
    a.kin_in = h->kin[h->cur];
    a.pos_out = h->pos[h->cur ^ 1];
    a.kin_out = h->kin[h->cur ^ 1];
    a.acc_out = h->acc;
    a.h = hs;
    a.hsq2 = 0.5 * (hs * hs);
    a.half_dt = 0.5 * hs;
    launch_pass<MODE_EULER, true>(h, a);
    // Time = Time + NewTimeStep; MoveWalls(Time)  (src/MDEM.f90:275-276)
    h->time = h->time + hs;
    enqueue_move_walls(h, h->time);
    after_eval(h, true, h->cur);
    h->cur ^= 1;
    enqueue_rebuild(h, h->cur, h->rb + (h->eval & 1), 0);
  }
  // PBC after the sub-steps (src/MDEM.f90:279), then the cell state of the
  // wrapped positions against the last sub-step's
  enqueue_refcell(h, h->cur, 1);
  HIPCHK(h, hipGetLastError());
  h->rc_ready = true;
  h->acc_in_ring = false;
  return check_errors(h);
}

}  // extern "C"

// Arguments of the two passes of one MoveParticles step (src/MDEM.f90:285-369),
// s = index of the step within the current call; advances the host-side
// bookkeeping (evaluation count, ipc rotation) as the launches will.
static void step_args(dem_handle_t h, int64_t s, bool first, PassArgs& a, PassArgs& b) {
  const size_t N = (size_t)h->np;
  // pass 1: forces at x_n (Update), AB predictor, ring store
  a = base_args(h);
  a.pos_in = h->pos[h->cur];
  a.kin_in = h->kin[h->cur];
  a.pos_out = h->pos[h->cur ^ 1];
  a.kin_out = h->kin[h->cur ^ 1];
  a.acc_out = h->acc;
  a.coef0 = h->ab0;
  for (int k = 0; k < h->nprev; ++k) {
    a.slot_rd[k] = h->ring + (size_t)(h->ipc[k] - 1) * 9 * N;
    a.coef[k] = h->ab[k];
  }
  a.slot_wr = h->nprev > 0 ? h->ring + (size_t)(h->ipc[h->nprev - 1] - 1) * 9 * N : nullptr;
  a.check_pred = first ? 1 : 0;
  after_eval(h, true, h->cur);
  if (h->nprev > 0) {
    const int last = h->ipc[h->nprev - 1];
    for (int k = h->nprev - 1; k > 0; --k) h->ipc[k] = h->ipc[k - 1];
    h->ipc[0] = last;
  }
  // pass 2: forces at the predicted state, AM corrector, look-ahead
  b = base_args(h);
  b.pos_in = h->pos[h->cur ^ 1];
  b.kin_in = h->kin[h->cur ^ 1];
  b.pos = h->pos[h->cur];
  b.kin = h->kin[h->cur];
  b.acc_out = h->keep_pred_acc ? h->acc_p : nullptr;
  b.coef0 = h->am0;
  for (int k = 0; k < h->nprev; ++k) {
    b.slot_rd[k] = h->ring + (size_t)(h->ipc[k] - 1) * 9 * N;
    b.coef[k] = h->am[k];
    b.ab[k] = h->ab[k];
  }
  b.ab0 = h->ab0;
  b.step_idx = (int)s;
  after_eval(h, false, h->cur ^ 1);
  h->acc_in_ring = h->nprev > 0;
}

// Undo the host-side bookkeeping of the steps enqueued after the one in
// which particles escaped: their kernels returned at once.
static void rollback_steps(dem_handle_t h, int64_t done, int64_t enqueued) {
  for (int64_t k = done; k < enqueued; ++k) {
    if (h->nprev > 0) {
      const int first_slot = h->ipc[0];
      for (int q = 0; q + 1 < h->nprev; ++q) h->ipc[q] = h->ipc[q + 1];
      h->ipc[h->nprev - 1] = first_slot;
    }
    h->eval -= 2;
    h->steps_since_resort -= 1;
  }
  h->last_update_eval = h->eval - 2;
}

extern "C" {

int dem_step(dem_handle_t h, int64_t nsteps, int64_t first_step, int64_t* steps_done,
             int32_t* escaped) {
  if (steps_done) *steps_done = 0;
  if (escaped) *escaped = 0;
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (nsteps <= 0) return DEM_OK;
  if (h->in_batch) return fail(h, DEM_ERR_STATE, "the system is stepped by a batch");
  if (h->np == 0) {
    // no particles: MoveParticles still turns the ring, sets Time and moves
    // the walls (src/MDEM.f90:329-337)
    for (int64_t s = 0; s < nsteps && h->nprev > 0; ++s) {
      const int last = h->ipc[h->nprev - 1];
      for (int k = h->nprev - 1; k > 0; --k) h->ipc[k] = h->ipc[k - 1];
      h->ipc[0] = last;
    }
    h->eval += 2 * nsteps;
    h->time = (double)(first_step + nsteps - 1) * h->dt;
    enqueue_move_walls(h, h->time);
    HIPCHK(h, hipGetLastError());
    h->next_step = first_step + nsteps;
    h->steps += nsteps;
    if (steps_done) *steps_done = nsteps;
    return DEM_OK;
  }
  hipStream_t st = h->stream;
  {
    const int rc = sync_halo(h, shape_staged(h->shape), st);
    if (rc) return rc;
  }
  {
    int big = 0x7fffffff;
    HIPCHK(h, hipMemcpyAsync(&h->ctl->stop_step, &big, sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(h, hipMemsetAsync(&h->ctl->escaped, 0, sizeof(int), st));
    HIPCHK(h, hipMemsetAsync(&h->ctl->esc_pending, 0, sizeof(int), st));
    HIPCHK(h, hipMemsetAsync(&h->ctl->pairs_upd_sum, 0, 2 * sizeof(unsigned long long), st));
  }
  if (!h->rc_ready) {
    enqueue_refcell(h, h->cur);
    h->rc_ready = true;
  }
  if (h->timing) h->timer.ensure((size_t)6 * nsteps + 2);
  size_t ev = 0;
  double t1 = 0, t2 = 0, tr = 0;
  bool first = true;
  bool have_psum = false;  // the previous step's corrector formed this predictor's sums
  int64_t enqueued = 0;
  for (int64_t s = 0; s < nsteps; ++s) {
    if (h->resort_interval > 0 && h->steps_since_resort >= h->resort_interval) {
      // a resort moves every particle array; never after an escape
      Ctl c0;
      int rc = read_ctl(h, &c0);
      if (rc) return rc;
      if (c0.escaped > 0) break;
      rc = resort(h);
      if (rc) return rc;
      have_psum = false;  // the sums were not permuted with the particles
    }
    PassArgs a, b;
    step_args(h, s, first, a, b);
    if (h->nprev > 0) {
      a.psum_rd = have_psum ? h->psum : nullptr;
      b.psum_wr = h->psum;
      have_psum = true;
    }
    if (h->timing) (void)hipEventRecord(h->timer.ev[ev++], st);
    launch_pass<MODE_PRED, true>(h, a);
    if (h->timing) (void)hipEventRecord(h->timer.ev[ev++], st);
    // Time = step*TimeStep; MoveWalls(Time)  (src/MDEM.f90:333-337)
    enqueue_move_walls(h, (double)(first_step + s) * h->dt);
    if (h->timing) (void)hipEventRecord(h->timer.ev[ev++], st);
    launch_pass<MODE_CORR, false>(h, b);
    if (h->timing) (void)hipEventRecord(h->timer.ev[ev++], st);
    // neighbour list, when pass 2 asked for it
    if (h->timing) (void)hipEventRecord(h->timer.ev[ev++], st);
    enqueue_rebuild(h, h->cur, h->rb + (h->eval & 1), 0);
    if (h->timing) (void)hipEventRecord(h->timer.ev[ev++], st);
    h->steps_since_resort += 1;
    first = false;
    enqueued = s + 1;
  }
  HIPCHK(h, hipGetLastError());
  Ctl c;
  int rc = read_ctl(h, &c);
  if (rc) return rc;
  if (h->timing) {
    float ms;
    for (size_t k = 0; k + 5 < ev; k += 6) {
      if (hipEventElapsedTime(&ms, h->timer.ev[k], h->timer.ev[k + 1]) == hipSuccess) t1 += ms;
      if (hipEventElapsedTime(&ms, h->timer.ev[k + 2], h->timer.ev[k + 3]) == hipSuccess) t2 += ms;
      if (hipEventElapsedTime(&ms, h->timer.ev[k + 4], h->timer.ev[k + 5]) == hipSuccess) tr += ms;
    }
    const double n = (double)nsteps;
    h->ms_pass1 = t1 / n;
    h->ms_pass2 = t2 / n;
    h->ms_rebuild = tr / n;
    if (ev >= 2 && hipEventElapsedTime(&ms, h->timer.ev[0], h->timer.ev[ev - 1]) == hipSuccess)
      h->ms_step = ms / n;
  }
  int64_t done = enqueued;
  if (c.escaped > 0) {
    // The kernels of the steps after the escape returned at once: undo their
    // host-side bookkeeping (ring rotation, evaluation count).
    done = (int64_t)c.stop_step + 1;
    rollback_steps(h, done, enqueued);
  }
  h->steps += done;
  if (done > 0) h->time = (double)(first_step + done - 1) * h->dt;
  h->next_step = first_step + done;
  if (steps_done) *steps_done = done;
  if (escaped) *escaped = c.escaped;
  rc = check_errors(h);
  if (rc) return rc;
  if (c.escaped > 0)
    return fail(h, DEM_ERR_ESCAPE, "%d particles left the box in step %lld of the call",
                c.escaped, (long long)done);
  return DEM_OK;
}

int dem_autotune(dem_handle_t h, int64_t steps_per_shape, int64_t first_step, dem_tune_t* out) {
  if (!h || !out) return DEM_ERR_ARG;
  memset(out, 0, sizeof *out);
  if (steps_per_shape < 1) return fail(h, DEM_ERR_ARG, "steps_per_shape must be at least 1");
  if (!h->has_state || h->np == 0) return fail(h, DEM_ERR_STATE, "upload a non-empty state first");
  const bool timing = h->timing;
  const int keep = h->shape;
  const int map0 = h->block_map;
  int64_t next = first_step;
  // one segment: an untimed step if warm (the first launch of a shape loads
  // its code), then nsteps steps timed with HIP events in shape c and block
  // order m; *ms = their summed time (both passes)
  auto seg = [&](int c, int m, int64_t nsteps, bool warm, double* ms) -> int {
    h->shape = c;
    h->block_map = m;
    *ms = 0.0;
    int64_t done = 0;
    if (warm) {
      h->timing = false;
      const int rc = dem_step(h, 1, next, &done, nullptr);
      out->steps += done;
      next += done;
      if (rc != DEM_OK) return rc;
    }
    if (nsteps <= 0) return DEM_OK;
    h->timing = true;
    const int rc = dem_step(h, nsteps, next, &done, nullptr);
    out->steps += done;
    next += done;
    if (rc == DEM_OK) *ms = (h->ms_pass1 + h->ms_pass2) * (double)nsteps;
    return rc;
  };
  auto undo = [&](int rc) {
    h->timing = timing;
    h->shape = keep;
    h->block_map = map0;
    return rc;
  };
  // every shape in a forward sweep, then in a backward sweep, half the steps
  // each: a trend in the system (contacts forming or relaxing) weighs on
  // every shape alike
  const int64_t h1 = (steps_per_shape + 1) / 2, h2 = steps_per_shape - h1;
  double t[DEM_NSHAPES], x;
  for (int c = 0; c < DEM_NSHAPES; ++c) {
    const int rc = seg(c, map0, h1, true, &t[c]);
    if (rc != DEM_OK) return undo(rc);
  }
  for (int c = DEM_NSHAPES - 1; c >= 0; --c) {
    const int rc = seg(c, map0, h2, false, &x);
    if (rc != DEM_OK) return undo(rc);
    t[c] += x;
  }
  int best = 0;
  for (int c = 0; c < DEM_NSHAPES; ++c) {
    out->ms[c] = t[c] / (double)steps_per_shape;
    if (out->ms[c] < out->ms[best]) best = c;
  }
  shape_table(out->block, out->waves, out->front);
  // the two block orders in the best shape, timed in the order A B B A
  double tm[2] = {0.0, 0.0};
  const struct {
    int map;
    int64_t nsteps;
    bool warm;
  } abba[4] = {{map0, h1, false}, {1 - map0, h1, true}, {1 - map0, h2, false}, {map0, h2, false}};
  for (const auto& q : abba) {
    const int rc = seg(best, q.map, q.nsteps, q.warm, &x);
    if (rc != DEM_OK) return undo(rc);
    tm[q.map] += x;
  }
  out->ms_map[0] = tm[0] / (double)steps_per_shape;
  out->ms_map[1] = tm[1] / (double)steps_per_shape;
  h->shape = best;
  h->block_map = (out->ms_map[1 - map0] < out->ms_map[map0]) ? 1 - map0 : map0;
  h->timing = timing;
  out->nshapes = DEM_NSHAPES;
  out->chosen = best;
  out->block_map = h->block_map;
  return DEM_OK;
}

int dem_synchronize(dem_handle_t h) {
  if (!h) return DEM_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DEM_OK;
}

int dem_stats(dem_handle_t h, dem_stats_t* out) {
  if (!h || !out) return DEM_ERR_ARG;
  memset(out, 0, sizeof *out);
  out->np = h->np;
  for (int d = 0; d < 3; ++d) {
    out->ref_ncell[d] = h->sc.rn[d];
    out->nb_ncell[d] = h->sc.gn[d];
  }
  out->time = h->time;
  out->next_step = h->next_step;
  if (!h->has_state) return DEM_OK;
  Ctl c;
  int rc = read_ctl(h, &c);
  if (rc) return rc;
  const int q = (int)((h->eval + 1) & 1);
  out->pair_contacts = (int64_t)c.pairs[q];
  out->wall_contacts = (int64_t)c.wallc[q];
  if (h->last_update_eval >= 0) out->stored_pairs = (int64_t)(c.stored[h->last_update_eval & 1] / 2);
  out->ncol = (int64_t)c.ncol;
  out->evaluations = h->eval;
  out->steps = h->steps;
  out->rebuilds = (int64_t)c.rebuilds;
  out->neighbour_pairs = (int64_t)c.nbpairs;
  out->stale_dropped = (int64_t)c.stale_dropped + h->uploaded_unlisted;
  if (h->np > 0) {
    const int nb = nblocks(h->np);
    hipLaunchKernelGGL(k_energy, dim3(nb), dim3(DEM_BLOCK), 0, h->stream, h->sc, h->pos[h->cur],
                       (const Kin*)h->kin[h->cur], h->d_epart);
    std::vector<double> part(nb);
    HIPCHK(h, hipMemcpyAsync(part.data(), h->d_epart, nb * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double e = 0.0;
    for (double x : part) e += x;
    out->kinetic_energy = 0.5 * e;
  }
  out->ms_pass1 = h->ms_pass1;
  out->ms_pass2 = h->ms_pass2;
  out->ms_rebuild = h->ms_rebuild;
  out->ms_step = h->ms_step;
  out->mean_pair_contacts = c.upd_evals ? (double)c.pairs_upd_sum / (double)c.upd_evals : 0.0;
  out->max_halo = c.max_halo;
  out->halo_overflow = c.halo_over;
  out->device_bytes = (int64_t)(h->state_bytes + const_bytes(h->sc.nc, h->sc.nw));
  return DEM_OK;
}

}  // extern "C"

// ---- observables -------------------------------------------------------------------

// The four observable kernels on the handle's stream with the partial sums in
// `scr` (obs_scratch_doubles(nl) doubles), then the results into out/profile.
static int run_observables(dem_handle_t h, int nl, double* scr, dem_observables_t* out,
                           double* profile) {
  memset(out, 0, sizeof *out);
  if (h->in_batch) HIPCHK(h, hipDeviceSynchronize());
  hipStream_t st = h->stream;
  ObsArgs o;
  o.pos = h->pos[h->cur];
  o.kin = h->kin[h->cur];
  o.E = h->es;
  o.nl = nl;
  o.nblk = std::max(1, std::min(DEM_OBS_BLOCKS, nblocks(h->np)));
  o.part = scr;
  o.lpart = scr + (size_t)DEM_OBS_BLOCKS * DEM_OBS_NPART;
  o.out = o.lpart + (size_t)DEM_OBS_BLOCKS * (size_t)nl * 4;
  hipLaunchKernelGGL(k_obs_partial, dim3(o.nblk), dim3(DEM_BLOCK), 0, st, h->sc, o);
  hipLaunchKernelGGL(k_obs_finish, dim3(1), dim3(DEM_BLOCK), 0, st, h->sc, o);
  hipLaunchKernelGGL(k_obs_bulk, dim3(o.nblk), dim3(DEM_BLOCK), 0, st, h->sc, o);
  hipLaunchKernelGGL(k_obs_phi, dim3(1), dim3(DEM_BLOCK), 0, st, h->sc, o);
  HIPCHK(h, hipGetLastError());
  double r[DEM_OBS_NOUT];
  HIPCHK(h, hipMemcpyAsync(r, o.out, sizeof r, hipMemcpyDeviceToHost, st));
  if (nl > 0 && profile)
    HIPCHK(h, hipMemcpyAsync(profile, o.out + DEM_OBS_NOUT, 4 * (size_t)nl * sizeof(double),
                             hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  out->kinetic_energy = r[OBS_KE];
  out->potential_energy = r[OBS_PE];
  out->pair_contacts = (int64_t)r[OBS_PAIRS];
  out->wall_contacts = (int64_t)r[OBS_WALLS];
  out->coordination = h->np > 0 ? 2.0 * r[OBS_PAIRS] / (double)h->np : 0.0;
  out->packing_fraction = r[OBS_PHI];
  out->region[0] = r[OBS_ZLO];
  out->region[1] = r[OBS_ZHI];
  out->solid_volume = r[OBS_VOL];
  out->top = r[OBS_TOP];
  return DEM_OK;
}

extern "C" {

int dem_observables(dem_handle_t h, int32_t nlayers, dem_observables_t* out, double* profile) {
  if (!h || !out) return DEM_ERR_ARG;
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (nlayers < 0 || nlayers > 65536)
    return fail(h, DEM_ERR_ARG, "nlayers must be 0..65536");
  if (nlayers > 0 && !profile) return fail(h, DEM_ERR_ARG, "null profile");
  const size_t need = obs_scratch_doubles(nlayers);
  if (h->obs_cap < need) {
    dfree(h->d_obs);
    h->obs_cap = 0;
    HIPCHK(h, dalloc(&h->d_obs, need));
    h->obs_cap = need;
  }
  return run_observables(h, nlayers, h->d_obs, out, profile);
}

int dem_observables_noalloc(dem_handle_t h, int32_t nlayers, dem_observables_t* out,
                            double* profile) {
  if (!h || !out) return DEM_ERR_ARG;
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (nlayers < 0 || nlayers > DEM_OBS_MAX_LAYERS)
    return fail(h, DEM_ERR_CAPACITY,
                "nlayers=%d: the allocation-free observables hold at most %d layers "
                "(dem_observables takes more)", nlayers, DEM_OBS_MAX_LAYERS);
  if (nlayers > 0 && !profile) return fail(h, DEM_ERR_ARG, "null profile");
  double* scr = g_scratch_ptr[h->device];
  if (!scr) return fail(h, DEM_ERR_STATE, "no scratch address for device %d", h->device);
  std::lock_guard<std::mutex> lk(g_obs_noalloc_mu);
  return run_observables(h, nlayers, scr, out, profile);
}

}  // extern "C"

// ---- particle number changes ---------------------------------------------------
//
// UpdateParticleList and PourParticle change the particle count between
// steps, rarely.  They run on the host, on the reference's own structures
// rebuilt from the device -- particles, ABM ring, cell lists in list order,
// contact lists -- and the result goes back with the cell state of the last
// evaluation kept (the next evaluation then transfers, as
// RedistributeParticles3D does, the particles whose positions left their
// cell).

namespace {

struct HEntry {
  int idx;
  double v1[3], v2[3];
};

struct HostSys {
  int np = 0;
  std::vector<int32_t> comp;                // 1-based
  std::vector<double> x, v, w;              // [np][3]
  std::vector<double> ring;                 // [np][27]: ring slot s at 9*s
  std::vector<double> pred_w;               // Predicted%Index - 1, by id
  std::vector<int> cell;                    // 0-based cell of the last evaluation
  std::vector<int> head, next;              // cell lists (head-pushed order)
  std::vector<std::vector<HEntry>> pairs;   // MainParticle(:)%FirstPartner
  std::vector<std::vector<HEntry>> walls;   // MainParticle(:)%FirstWall
};

}  // namespace

static int host_download(dem_handle_t h, HostSys& hs) {
  const int np = h->np;
  const size_t N = (size_t)np;
  hs.np = np;
  hs.comp.assign(N, 0);
  hs.x.assign(3 * N, 0.0);
  hs.v.assign(3 * N, 0.0);
  hs.w.assign(3 * N, 0.0);
  hs.ring.assign(27 * N, 0.0);
  hs.cell.assign(N, 0);
  const int ncell = h->sc.rn[0] * h->sc.rn[1] * h->sc.rn[2];
  hs.head.assign((size_t)ncell, -1);
  hs.next.assign(N, -1);
  const size_t nl = (size_t)std::max(np, h->max_np);
  hs.pairs.assign(nl, {});
  hs.walls.assign(nl, {});
  std::vector<uint64_t> keys, wkeys;
  std::vector<double> ft, vt, wft, wvt;
  if (np > 0) {
    std::vector<double> xs(3 * N), vs(3 * N), ws(3 * N), ring((size_t)std::max(h->nprev, 1) * 9 * N);
    std::vector<int32_t> ipc(3), cells(N), rank(N);
    int rc = dem_download_state(h, np, hs.comp.data(), xs.data(), vs.data(), ws.data(), nullptr,
                                nullptr, ring.data(), ipc.data());
    if (!rc) rc = dem_download_cells(h, np, cells.data());
    if (!rc) rc = dem_download_cell_rank(h, np, rank.data());
    if (rc) return rc;
    for (size_t i = 0; i < N; ++i) {
      for (int d = 0; d < 3; ++d) {
        hs.x[3 * i + d] = xs[d * N + i];
        hs.v[3 * i + d] = vs[d * N + i];
        hs.w[3 * i + d] = ws[d * N + i];
      }
      for (int k = 0; k < h->nprev; ++k)
        for (int q = 0; q < 9; ++q) hs.ring[27 * i + 9 * k + q] = ring[((size_t)k * 9 + q) * N + i];
      hs.cell[i] = cells[i] - 1;
    }
    // lists: within a cell by rank, highest first (head)
    std::vector<int> order(N);
    for (size_t i = 0; i < N; ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
      if (hs.cell[a] != hs.cell[b]) return hs.cell[a] < hs.cell[b];
      return rank[a] < rank[b];
    });
    for (int q : order) {  // ascending rank: each push lands at the head
      hs.next[q] = hs.head[hs.cell[q]];
      hs.head[hs.cell[q]] = q;
    }
  }
  int64_t npairs = 0, nwall = 0;
  int rc = dem_download_contacts(h, 0, &npairs, nullptr, nullptr, nullptr, 0, &nwall, nullptr,
                                 nullptr, nullptr);
  if (rc) return rc;
  keys.resize((size_t)npairs);
  ft.resize(3 * (size_t)npairs);
  vt.resize(3 * (size_t)npairs);
  wkeys.resize((size_t)nwall);
  wft.resize(3 * (size_t)nwall);
  wvt.resize(3 * (size_t)nwall);
  rc = dem_download_contacts(h, npairs, &npairs, keys.data(), ft.data(), vt.data(), nwall, &nwall,
                             wkeys.data(), wft.data(), wvt.data());
  if (rc) return rc;
  for (int64_t k = 0; k < npairs; ++k) {
    const size_t a = (size_t)(keys[k] >> 32) - 1;
    HEntry e;
    e.idx = (int)(keys[k] & 0xffffffffu) - 1;
    memcpy(e.v1, &ft[3 * k], sizeof e.v1);
    memcpy(e.v2, &vt[3 * k], sizeof e.v2);
    if (a >= hs.pairs.size()) hs.pairs.resize(a + 1);
    hs.pairs[a].push_back(e);
  }
  for (int64_t k = 0; k < nwall; ++k) {
    const size_t a = (size_t)(wkeys[k] >> 32) - 1;
    HEntry e;
    e.idx = (int)(wkeys[k] & 0xffffffffu) - 1;
    memcpy(e.v1, &wft[3 * k], sizeof e.v1);
    memcpy(e.v2, &wvt[3 * k], sizeof e.v2);
    hs.walls[a].push_back(e);
  }
  hs.pred_w.assign(N, 0.0);
  for (size_t i = 0; i < N; ++i)
    hs.pred_w[i] = h->pred_w_h.empty() ? (double)(hs.comp[i] - 1) : h->pred_w_h[i];
  return DEM_OK;
}

static int host_upload(dem_handle_t h, const HostSys& hs, int max_np) {
  Ctl c0;
  int rc = read_ctl(h, &c0);
  if (rc) return rc;
  const int np = hs.np;
  const size_t N = (size_t)np;
  std::vector<double> xs(3 * N), vs(3 * N), ws(3 * N), ring((size_t)std::max(h->nprev, 1) * 9 * N);
  std::vector<int32_t> rank(N, 0);
  std::vector<int2> cr(N);
  for (size_t i = 0; i < N; ++i) {
    for (int d = 0; d < 3; ++d) {
      xs[d * N + i] = hs.x[3 * i + d];
      vs[d * N + i] = hs.v[3 * i + d];
      ws[d * N + i] = hs.w[3 * i + d];
    }
    for (int k = 0; k < h->nprev; ++k)
      for (int q = 0; q < 9; ++q) ring[((size_t)k * 9 + q) * N + i] = hs.ring[27 * i + 9 * k + q];
  }
  for (size_t c = 0; c < hs.head.size(); ++c) {
    int n = 0;
    for (int q = hs.head[c]; q >= 0; q = hs.next[q]) ++n;
    for (int q = hs.head[c]; q >= 0; q = hs.next[q]) rank[q] = --n;
  }
  for (size_t i = 0; i < N; ++i) cr[i] = make_int2(hs.cell[i], rank[i]);
  int32_t ipc[3] = {h->ipc[0], h->ipc[1], h->ipc[2]};
  const double time = h->time;
  const int64_t next_step = h->next_step;
  bool stale = false;
  for (size_t i = 0; i < N; ++i) stale = stale || hs.pred_w[i] != (double)(hs.comp[i] - 1);
  h->pending_cut_all = stale ? 1 : 0;  // the neighbour lists must cover the other radius
  rc = dem_upload_state(h, np, hs.comp.data(), xs.data(), vs.data(), ws.data(), nullptr, nullptr,
                        h->nprev > 0 ? ring.data() : nullptr, h->nprev > 0 ? ipc : nullptr,
                        rank.data());
  if (rc) return rc;
  h->time = time;
  h->next_step = next_step;
  h->max_np = max_np;
  hipStream_t st = h->stream;
  // contacts: stored pairs between particles 1..NP go to the device, the
  // others stay on the host until the particle count covers them
  std::vector<uint64_t> keys, wkeys;
  std::vector<double> ft, vt, wft, wvt;
  std::vector<dem_system::Orphan> orphans;
  for (size_t a = 0; a < hs.pairs.size(); ++a) {
    std::vector<HEntry> l = hs.pairs[a];
    std::stable_sort(l.begin(), l.end(), [](const HEntry& p, const HEntry& q) { return p.idx < q.idx; });
    for (const HEntry& e : l) {
      const uint64_t key = ((uint64_t)(a + 1) << 32) | (uint64_t)(e.idx + 1);
      if ((int)a < np && e.idx < np) {
        keys.push_back(key);
        ft.insert(ft.end(), e.v1, e.v1 + 3);
        vt.insert(vt.end(), e.v2, e.v2 + 3);
      } else {
        dem_system::Orphan o;
        o.key = key;
        memcpy(o.ft, e.v1, sizeof o.ft);
        memcpy(o.vt, e.v2, sizeof o.vt);
        orphans.push_back(o);
      }
    }
  }
  for (size_t a = 0; a < hs.walls.size() && (int)a < np; ++a)
    for (const HEntry& e : hs.walls[a]) {
      wkeys.push_back(((uint64_t)(a + 1) << 32) | (uint64_t)(e.idx + 1));
      wft.insert(wft.end(), e.v1, e.v1 + 3);
      wvt.insert(wvt.end(), e.v2, e.v2 + 3);
    }
  if (np > 0) {
    rc = dem_upload_contacts(h, (int64_t)keys.size(), keys.data(), ft.data(), vt.data(),
                             (int64_t)wkeys.size(), wkeys.data(), wft.data(), wvt.data());
    if (rc) return rc;
  }
  h->orphans = orphans;
  if (np > 0) {
    // the cell state of the last evaluation, lists in the given order
    int2* d_cr = nullptr;
    HIPCHK(h, hipMalloc((void**)&d_cr, N * sizeof(int2)));
    hipError_t e = hipMemcpyAsync(d_cr, cr.data(), N * sizeof(int2), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_set_ref_cells, dim3(nblocks(np)), dim3(DEM_BLOCK), 0, st, h->sc,
                         h->kin[h->cur], (const int2*)d_cr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_cr);
    if (e != hipSuccess) return fail(h, DEM_ERR_DEVICE, "cell state upload: %s", hipGetErrorString(e));
    h->last_eval_buf = h->cur;
    h->rc_ready = false;
  }
  // Predicted%Index where it differs from the particle's component
  h->pred_w_h.clear();
  if (stale) {
    h->pred_w_h = hs.pred_w;
    if (h->pred_w_cap < N) {
      dfree(h->d_pred_w);
      HIPCHK(h, dalloc(&h->d_pred_w, N));
      h->pred_w_cap = N;
    }
    HIPCHK(h, hipMemcpy(h->d_pred_w, h->pred_w_h.data(), N * sizeof(double), hipMemcpyHostToDevice));
  }
  // counters that run across the whole simulation
  HIPCHK(h, hipMemcpyAsync(&h->ctl->ncol, &c0.ncol, 2 * sizeof(unsigned long long),
                           hipMemcpyHostToDevice, st));
  HIPCHK(h, hipMemcpyAsync(&h->ctl->stale_dropped, &c0.stale_dropped, sizeof(unsigned long long),
                           hipMemcpyHostToDevice, st));
  HIPCHK(h, hipStreamSynchronize(st));
  return check_errors(h);
}

// CurrentCell3D (src/MCells.f90:498-513), 0-based.
static int host_cell(const SysConst& s, const double* r) {
  const RefCell c = ref_cell(s, r[0], r[1], r[2]);
  return c.lin;
}

// The Monte Carlo neighbour list of a cell: its own list, then the lists of
// the (up to 26) cells around it in the head-pushed order of AddNeighbour
// (HinderNeighbourCellsList, NeighbourCells(iCell, MonteCarlo, ...),
// AllNeighbourParticles: src/MCells.f90:340-374,424-476,833-860,1056-1074).
static void host_mc_neighbours(const SysConst& s, const HostSys& hs, int icell,
                               std::vector<int>& out) {
  const int nx = s.rn[0], ny = s.rn[1], nz = s.rn[2];
  const int iz = icell / (nx * ny);
  const int iy = icell / nx - iz * ny;
  const int ix = icell - (iy + iz * ny) * nx;
  auto range = [](int i, int m, int pbc, int& lo, int& hi) {
    lo = -1;
    hi = 1;
    if (!pbc) {
      if (i == 0) lo = 0;
      if (i == m - 1) hi = 0;
    }
  };
  auto wrap = [](int j, int m, int pbc) {
    if (pbc) {
      if (j == m) j = 0;
      if (j == -1) j = m - 1;
    }
    return j;
  };
  int xi, xf, yi, yf, zi, zf;
  range(ix, nx, s.pbc[0], xi, xf);
  range(iy, ny, s.pbc[1], yi, yf);
  range(iz, nz, s.pbc[2], zi, zf);
  std::vector<int> cells;
  for (int kz = zi; kz <= zf; ++kz) {
    const int jz = wrap(iz + kz, nz, s.pbc[2]);
    for (int ky = yi; ky <= yf; ++ky) {
      const int jy = wrap(iy + ky, ny, s.pbc[1]);
      for (int kx = xi; kx <= xf; ++kx) {
        if (kx + nx * (ky + ny * kz) == 0) continue;
        const int jx = wrap(ix + kx, nx, s.pbc[0]);
        cells.push_back(jx + nx * (jy + ny * jz));
      }
    }
  }
  out.clear();
  for (int q = hs.head[icell]; q >= 0; q = hs.next[q]) out.push_back(q);
  for (size_t k = cells.size(); k-- > 0;)
    for (int q = hs.head[cells[k]]; q >= 0; q = hs.next[q]) out.push_back(q);
}

// ResizeVariables (src/MDEM.f90:566-608): MainParticle keeps the lists of
// particles 1..NP, Predicted%Index is reset to the components.
static void host_resize(HostSys& hs, int& max_np, int m) {
  for (size_t a = (size_t)hs.np; a < hs.pairs.size(); ++a) {
    hs.pairs[a].clear();
    hs.walls[a].clear();
  }
  max_np = m;
  if (hs.pairs.size() < (size_t)m) {
    hs.pairs.resize((size_t)m);
    hs.walls.resize((size_t)m);
  }
  for (int i = 0; i < hs.np; ++i) hs.pred_w[i] = (double)(hs.comp[i] - 1);
}

extern "C" {

// ranf (src/MUtil.f90:16-44).
double dem_ranf(dem_rng_t* r) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  if (r->dummy <= 0 || r->iiy < 0) {
    r->aam = (double)(std::nextafter(1.0f, 0.0f) / (float)im);
    const int32_t a = r->dummy < 0 ? -r->dummy : r->dummy;
    r->iiy = (888889999 ^ a) | 1;
    r->iix = 777755555 ^ a;
    r->dummy = a + 1;
  }
  uint32_t x = (uint32_t)r->iix;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  r->iix = (int32_t)x;
  r->iik = r->iiy / iq;
  r->iiy = ia * (r->iiy - r->iik * iq) - ir * r->iik;
  if (r->iiy < 0) r->iiy += im;
  return r->aam * (double)((im & (r->iix ^ r->iiy)) | 1);
}

// UpdateParticleList (src/MDEM.f90:612-671).
int dem_remove_escaped(dem_handle_t h, int32_t cap, int32_t* removed, int32_t* nremoved) {
  if (nremoved) *nremoved = 0;
  if (!h || !h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (h->in_batch) return fail(h, DEM_ERR_STATE, "the system is stepped by a batch");
  HostSys hs;
  int rc = host_download(h, hs);
  if (rc) return rc;
  const SysConst& s = h->sc;
  int max_np = h->max_np;
  int count = 0;
  std::vector<int> nb;
  for (int i = hs.np - 1; i >= 0; --i) {
    const double* r = &hs.x[3 * (size_t)i];
    bool out = false;
    for (int d = 0; d < 3; ++d) out = out || (r[d] > s.phalf[d]) || (r[d] < -s.phalf[d]);
    if (!out) continue;
    const int icell = hs.cell[i];
    host_mc_neighbours(s, hs, icell, nb);
    // RemoveParticleFromCell (src/MCells.f90:702-724)
    {
      int* link = &hs.head[icell];
      while (*link != i) link = &hs.next[*link];
      *link = hs.next[i];
    }
    // RemoveParticleFromContactList (src/MContacts.f90:579-633)
    for (int j : nb) {
      auto& l = hs.pairs[j];
      if (l.empty()) continue;
      size_t pos = 0;
      while (l[pos].idx < i && pos + 1 < l.size()) ++pos;
      if (l[pos].idx == i) l.erase(l.begin() + (long)pos);
    }
    hs.pairs[i].clear();
    hs.walls[i].clear();
    const int last = hs.np - 1;
    if (i != last) {
      hs.comp[i] = hs.comp[last];
      for (int d = 0; d < 3; ++d) {
        hs.x[3 * (size_t)i + d] = hs.x[3 * (size_t)last + d];
        hs.v[3 * (size_t)i + d] = hs.v[3 * (size_t)last + d];
        hs.w[3 * (size_t)i + d] = hs.w[3 * (size_t)last + d];
      }
      std::copy(&hs.ring[27 * (size_t)last], &hs.ring[27 * (size_t)last] + 27, &hs.ring[27 * (size_t)i]);
      // ReplaceParticleIndexInCell: the entry keeps its place
      {
        int* link = &hs.head[hs.cell[last]];
        while (*link != last) link = &hs.next[*link];
        *link = i;
        hs.next[i] = hs.next[last];
      }
      hs.cell[i] = hs.cell[last];
      // ChangeIndexInContactList (src/MContacts.f90:173-252)
      std::vector<HEntry> first;
      for (int j : nb) {
        auto& l = hs.pairs[j];
        if (l.empty() || l.back().idx != last) continue;
        HEntry e = l.back();
        if (j < i) {
          if (l.size() == 1) {
            l[0].idx = i;
          } else {
            int point = -1;
            for (size_t k = 0; k + 1 < l.size(); ++k)
              if (l[k].idx < i) point = (int)k;
            l.pop_back();
            e.idx = i;
            l.insert(l.begin() + (point + 1), e);
          }
        } else {
          l.pop_back();
          e.idx = j;
          first.push_back(e);
        }
      }
      hs.pairs[i] = first;
      hs.walls[i] = hs.walls[last];
      hs.walls[last].clear();
    }
    hs.np -= 1;
    if (removed && count < cap) removed[count] = i + 1;
    ++count;
    if (max_np >= (int)(1.1 * (double)hs.np)) host_resize(hs, max_np, (int)(1.05 * (double)hs.np));
  }
  if (nremoved) *nremoved = count;
  if (count == 0) return DEM_OK;
  // the arrays by particle end at the new NP (Predicted%Index of the slots
  // that received a particle keeps the old component)
  const size_t N = (size_t)hs.np;
  hs.comp.resize(N);
  hs.x.resize(3 * N);
  hs.v.resize(3 * N);
  hs.w.resize(3 * N);
  hs.ring.resize(27 * N);
  hs.cell.resize(N);
  hs.next.resize(N);
  hs.pred_w.resize(N);
  return host_upload(h, hs, max_np);
}

// PourParticle and InsertParticle (src/MDEM.f90:412-562).  The reference
// tests the spacing with the integer parts of the separations (Rij is
// declared integer) against MinDijSq of the largest component, and the
// walls with the largest radius; the new particle falls at
// -1e-4*LengthScale/TimeScale.
int dem_pour(dem_handle_t h, const dem_pour_t* pr, dem_rng_t* rng, int32_t* inserted) {
  if (inserted) *inserted = 0;
  if (!h || !pr || !rng || !pr->target_number) return fail(h, DEM_ERR_ARG, "null argument");
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (h->in_batch) return fail(h, DEM_ERR_STATE, "the system is stepped by a batch");
  if (pr->shape != DEM_POUR_RECTANGLE && pr->shape != DEM_POUR_CIRCLE)
    return fail(h, DEM_ERR_ARG, "bad pouring shape %d", pr->shape);
  const SysConst& s = h->sc;
  const int nc = s.nc;
  HostSys hs;
  int rc = host_download(h, hs);
  if (rc) return rc;
  std::vector<int> number((size_t)nc, 0), notreached;
  for (int i = 0; i < hs.np; ++i) number[hs.comp[i] - 1] += 1;
  for (int c = 0; c < nc; ++c)
    if (number[c] < pr->target_number[c]) notreached.push_back(c + 1);
  if (notreached.empty()) return DEM_OK;
  const int nnr = (int)notreached.size();
  const int j0 = (int)((double)nnr * dem_ranf(rng)) + 1;
  const int index = notreached[j0 - 1];
  int ibig = 0;
  for (int c = 1; c < nc; ++c)
    if (h->radius_h[c] > h->radius_h[ibig]) ibig = c;
  const double ri = h->radius_h[ibig], di = 2.0 * ri;
  double p[3] = {0.0, 0.0, s.phalf[2] - ri};
  const double two_pi = 2.0 * 3.14159265358979323846;
  // MinDijSq of the largest component with each component
  std::vector<PairConst> pair((size_t)nc * nc);
  HIPCHK(h, hipMemcpy(pair.data(), h->d_pair, pair.size() * sizeof(PairConst), hipMemcpyDeviceToHost));
  bool possible = false;
  int attempt = 0;
  std::vector<int> nb;
  while (!possible && attempt < pr->max_attempts) {
    possible = true;
    ++attempt;
    if (pr->shape == DEM_POUR_RECTANGLE) {
      for (int d = 0; d < 2; ++d) p[d] = (pr->size[d] - di) * (dem_ranf(rng) - 0.5);
    } else {
      const double rad = 0.5 * dem_ranf(rng) * (pr->size[0] - di);
      const double angle = two_pi * dem_ranf(rng);
      p[0] = rad * std::sin(angle);
      p[1] = rad * std::cos(angle);
    }
    host_mc_neighbours(s, hs, host_cell(s, p), nb);
    for (size_t k = 0; possible && k < nb.size(); ++k) {
      const double* q = &hs.x[3 * (size_t)nb[k]];
      int32_t sum = 0;
      for (int d = 0; d < 3; ++d) {
        const int32_t r = (int32_t)(q[d] - p[d]);
        sum += r * r;
      }
      possible = (double)sum > pair[(size_t)ibig + (size_t)(hs.comp[nb[k]] - 1) * nc].mindijsq;
    }
    for (int jw = 0; possible && jw < s.nw; ++jw) {
      const WallConst& wc = h->wallc_h[(size_t)(index - 1) + (size_t)jw * nc];
      const WallHit hit = wall_overlap(s, h->walls_h[jw], mk3(p[0], p[1], p[2]), ri,
                                       wc.max_alpha_u, wc.min_aw_sq, wc.max_aw_sq);
      if (hit.ov < 0)
        return fail(h, DEM_ERR_UNSUPPORTED, "pouring position on the axis of cone wall %d", jw + 1);
      possible = hit.ov == 0;
    }
  }
  if (!possible) return DEM_OK;
  int max_np = h->max_np;
  if (hs.np == max_np) host_resize(hs, max_np, std::max(max_np + 1, (int)(1.05 * (double)max_np)));
  // InsertParticle (src/MDEM.f90:540-562)
  const int k = hs.np;
  hs.np += 1;
  hs.comp.push_back(index);
  for (int d = 0; d < 3; ++d) hs.x.push_back(p[d]);
  hs.v.push_back(0.0);
  hs.v.push_back(0.0);
  hs.v.push_back(-(1.e-4 * pr->length_scale / pr->time_scale));
  for (int d = 0; d < 3; ++d) hs.w.push_back(0.0);
  hs.ring.resize(27 * (size_t)hs.np, 0.0);
  hs.pred_w.push_back((double)(index - 1));
  const int c = host_cell(s, p);
  hs.cell.push_back(c);
  hs.next.push_back(hs.head[c]);
  hs.head[c] = k;
  if (hs.pairs.size() < (size_t)hs.np) {
    hs.pairs.resize((size_t)hs.np);
    hs.walls.resize((size_t)hs.np);
  }
  if (inserted) *inserted = 1;
  return host_upload(h, hs, max_np);
}

}  // extern "C"

// ---- batches ---------------------------------------------------------------

struct dem_batch {
  char err[512];
  std::vector<dem_handle_t> sys;
  hipStream_t stream = nullptr;
  int precision = -1;  // DEM_FP64 / DEM_MIXED of its systems
  static const int K = 4;  // pinned staging slots (a step's table is copied from one)
  BatchRec* h_rec[K] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev[K] = {nullptr, nullptr, nullptr, nullptr};
  bool ev_used[K] = {false, false, false, false};
  size_t rec_cap = 0;
  BatchRec* d_rec = nullptr;
  int2* d_map[MAP_KINDS] = {};
  size_t map_cap[MAP_KINDS] = {};
  int map_len[MAP_KINDS] = {};
  bool dirty = true;
  Ctl* d_ctl = nullptr;
  Ctl* h_ctl = nullptr;
  size_t ctl_cap = 0;
  bool timing = false;
  std::vector<hipEvent_t> tev;
  std::vector<double> step_ms;
  // Captured step sequences: gexec[0] replays one step, gexec[1]
  // graph_steps steps.  Valid while the table, the block maps and the
  // launch shapes they were captured with stay the same (gkey).
  bool use_graph = false;
  int graph_steps = 8;
  int shape = 0;  // launch shape of the force passes (DEM_SHAPE_*)
  int block_map = 0;  // 0: particle ranges per XCD, 1: dispatch order
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<long long> gkey;
  // Persistent stepping (k_batch_persistent): steps per cooperative launch,
  // the group of every block, the group sizes, barrier words, merge scratch
  // and the abort word.
  bool persistent = false;
  int persist_steps = 250;
  unsigned spin_limit = 1u << 24;
  int2* d_group = nullptr;
  size_t group_cap = 0;
  int* d_gsize = nullptr;
  GroupSync* d_gsync = nullptr;
  MergeScratch* d_scratch = nullptr;
  char* d_scratch_mem = nullptr;
  size_t pers_cap = 0;  // systems the per-system arrays hold
  int* d_abort = nullptr;
  int pers_grid = 0;
  std::vector<int> pers_key;  // pass blocks per system the groups were cut for
};

// Whether dem_batch_step runs the persistent launch (per-step timing needs
// the launch-per-phase path).
static bool batch_persistent_now(const dem_batch* b) { return b->persistent && !b->timing; }
// FRONT of the persistent kernel: the batch's launch shape's when it has
// DEM_BLOCK lanes, else the default shape's.
static int batch_persistent_front(const dem_batch* b);

static int bfail(dem_batch_t b, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b ? b->err : g_err, 512, fmt, ap);
  va_end(ap);
  return code;
}

#define BHIPCHK(b, expr)                                                                 \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return bfail((b), DEM_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// Arguments of the neighbour-list rebuild of system h around pos[b] (the
// kernels of enqueue_rebuild), gated on *gate.
static RebuildArgs rebuild_args(dem_handle_t h, int b, const int* gate) {
  RebuildArgs r;
  memset(&r, 0, sizeof r);
  r.pos = h->pos[b];
  r.kin = h->kin[b];
  r.xref = h->xref;
  r.key = h->key;
  r.cell_count = h->cell_count;
  r.cell_start = h->cell_start;
  r.cursor = h->cursor;
  r.cell_idx = h->cell_idx;
  r.tile_sum = h->tile_sum;
  r.slice_w = h->slice_w;
  r.d_es = h->d_es;
  r.es = h->es;
  r.gate = gate;
  r.ctl = h->ctl;
  r.gncell = h->gncell;
  r.ntiles = h->ntiles;
  r.nslices = h->nslices;
  r.ntiles_np = h->ntiles_np;
  return r;
}

// Table entry of system h at the start of a dem_batch_step call: constants,
// the step-invariant arguments of its passes and rebuild (those step_args
// and enqueue_rebuild give) and its bookkeeping; k_batch_prepare fills in
// the per-step arguments.
static void batch_rec_init(dem_handle_t h, BatchRec& r) {
  memset(&r, 0, sizeof r);
  r.s = h->sc;
  PassArgs a = base_args(h);
  a.pos_in = h->pos[h->cur];
  a.kin_in = h->kin[h->cur];
  a.pos_out = h->pos[h->cur ^ 1];
  a.kin_out = h->kin[h->cur ^ 1];
  a.acc_out = h->acc;
  a.coef0 = h->ab0;
  for (int k = 0; k < h->nprev; ++k) a.coef[k] = h->ab[k];
  PassArgs b = base_args(h);
  b.pos_in = h->pos[h->cur ^ 1];
  b.kin_in = h->kin[h->cur ^ 1];
  b.pos = h->pos[h->cur];
  b.kin = h->kin[h->cur];
  b.acc_out = h->keep_pred_acc ? h->acc_p : nullptr;
  b.coef0 = h->am0;
  for (int k = 0; k < h->nprev; ++k) {
    b.coef[k] = h->am[k];
    b.ab[k] = h->ab[k];
  }
  b.ab0 = h->ab0;
  r.p[0] = a;
  r.p[1] = b;
  r.r = rebuild_args(h, h->cur, h->rb);
  r.nblk = nblocks(h->np);
  r.pass_nblk = nblocks(h->np, h->batch && batch_persistent_now(h->batch)
                                   ? DEM_BLOCK
                                   : shape_block(h->batch ? h->batch->shape : 0));
  r.nprev = h->nprev;
  r.first = h->batch_first ? 1 : 0;
  r.step = 0;
  for (int k = 0; k < 3; ++k) r.ipc[k] = h->ipc[k];
  r.eval = h->eval;
  r.stamp_base = h->stamp_base;
  r.ring = h->ring;
  r.psum = h->psum;
  r.rb = h->rb;
  r.walls = h->d_walls;
  r.wm = h->d_wm;
  r.nw_move = h->nw_move;
  r.next_step = h->next_step;
  r.dt = h->dt;
}

// Host-side bookkeeping of n batched steps (what step_args does per step).
static void batch_advance_host(dem_handle_t h, int64_t n) {
  for (int64_t s = 0; s < n; ++s) {
    after_eval(h, true, h->cur);
    if (h->nprev > 0) {
      const int last = h->ipc[h->nprev - 1];
      for (int k = h->nprev - 1; k > 0; --k) h->ipc[k] = h->ipc[k - 1];
      h->ipc[0] = last;
    }
    after_eval(h, false, h->cur ^ 1);
  }
  h->acc_in_ring = h->nprev > 0;
  h->batch_first = false;
  h->steps_since_resort += n;
}

// Block maps of the current systems: particles (256 per block), cell keys
// (256 per block), cell-scan tiles, slice-scan tiles.
static int batch_maps(dem_batch_t b) {
  std::vector<int2> m[MAP_KINDS];
  for (size_t i = 0; i < b->sys.size(); ++i) {
    const dem_handle_t h = b->sys[i];
    const int cnt[MAP_KINDS] = {nblocks(h->np), nblocks((long long)h->gncell + 1), h->ntiles,
                                h->ntiles_np, nblocks(h->np, shape_block(b->shape))};
    for (int k = 0; k < MAP_KINDS; ++k)
      for (int j = 0; j < cnt[k]; ++j) m[k].push_back(make_int2((int)i, j));
  }
  for (int k = 0; k < MAP_KINDS; ++k) {
    if (m[k].size() > b->map_cap[k]) {
      if (b->d_map[k]) (void)hipFree(b->d_map[k]);
      b->d_map[k] = nullptr;
      BHIPCHK(b, hipMalloc((void**)&b->d_map[k], m[k].size() * sizeof(int2)));
      b->map_cap[k] = m[k].size();
    }
    if (!m[k].empty())
      BHIPCHK(b, hipMemcpyAsync(b->d_map[k], m[k].data(), m[k].size() * sizeof(int2),
                                hipMemcpyHostToDevice, b->stream));
    b->map_len[k] = (int)m[k].size();
  }
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  b->dirty = false;
  return DEM_OK;
}

static int batch_capacity(dem_batch_t b, size_t n) {
  if (n <= b->rec_cap) return DEM_OK;
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  for (int k = 0; k < dem_batch::K; ++k) {
    if (b->h_rec[k]) (void)hipHostFree(b->h_rec[k]);
    b->h_rec[k] = nullptr;
    BHIPCHK(b, hipHostMalloc((void**)&b->h_rec[k], n * sizeof(BatchRec), hipHostMallocDefault));
    b->ev_used[k] = false;
  }
  if (b->d_rec) (void)hipFree(b->d_rec);
  b->d_rec = nullptr;
  BHIPCHK(b, hipMalloc((void**)&b->d_rec, n * sizeof(BatchRec)));
  if (b->d_ctl) (void)hipFree(b->d_ctl);
  if (b->h_ctl) (void)hipHostFree(b->h_ctl);
  b->d_ctl = nullptr;
  b->h_ctl = nullptr;
  BHIPCHK(b, hipMalloc((void**)&b->d_ctl, n * sizeof(Ctl)));
  BHIPCHK(b, hipHostMalloc((void**)&b->h_ctl, n * sizeof(Ctl), hipHostMallocDefault));
  b->rec_cap = n;
  return DEM_OK;
}

// The batched pass kernel in launch shape `shape` over the pass-block map.
template <int MODE, bool UPDATE, bool MIXED, int PI, int S = 0>
static void launch_batch_shape(int shape, int nb, hipStream_t st, const BatchRec* rec,
                               const int2* map, int xcd) {
  if constexpr (S < DEM_NSHAPES) {
    using Sh = PassShape<S>;
    if (shape == S) {
      hipLaunchKernelGGL((k_pass_batch<MODE, UPDATE, MIXED, PI, Sh::block, Sh::waves, Sh::front>),
                         dim3(nb), dim3(Sh::block), 0, st, rec, map, xcd);
      return;
    }
    launch_batch_shape<MODE, UPDATE, MIXED, PI, S + 1>(shape, nb, st, rec, map, xcd);
  }
}

// The launches of one batched step (every phase is one launch for all the
// systems; the arguments live in the device table, so the launches are the
// same from step to step and can be replayed from a captured graph).
static void enqueue_batch_step(dem_batch_t b, int ns, bool mixed, bool moving) {
  hipStream_t st = b->stream;
  BatchRec* rec = b->d_rec;
  const int gp = b->map_len[MAP_PART], gc = b->map_len[MAP_CELL];
  const int gt = b->map_len[MAP_CTILE], gs = b->map_len[MAP_STILE];
  const int2* mp = b->d_map[MAP_PART];
  // pass 1, pass 2 (one launch each for all systems) and their counters;
  // the step's arguments were prepared by k_batch_prepare (first step of the
  // call) or by the previous step's flip
  const int gq = b->map_len[MAP_PASS];
  const int2* mq = b->d_map[MAP_PASS];
  if (mixed)
    launch_batch_shape<MODE_PRED, true, true, 0>(b->shape, gq, st, rec, mq,
                                                    b->block_map == 0);
  else
    launch_batch_shape<MODE_PRED, true, false, 0>(b->shape, gq, st, rec, mq,
                                                    b->block_map == 0);
  hipLaunchKernelGGL(k_pass_reduce_batch, dim3(ns), dim3(1024), 0, st, rec, 0, 1);
  if (moving) hipLaunchKernelGGL(k_move_walls_batch, dim3(ns), dim3(64), 0, st, rec);
  if (mixed)
    launch_batch_shape<MODE_CORR, false, true, 1>(b->shape, gq, st, rec, mq,
                                                    b->block_map == 0);
  else
    launch_batch_shape<MODE_CORR, false, false, 1>(b->shape, gq, st, rec, mq,
                                                    b->block_map == 0);
  hipLaunchKernelGGL(k_pass_reduce_batch, dim3(ns), dim3(1024), 0, st, rec, 1, 0);
  // neighbour lists of the systems whose pass 2 asked for one
  hipLaunchKernelGGL(k_bin_count_batch, dim3(gp), dim3(DEM_BLOCK), 0, st, rec, mp);
  hipLaunchKernelGGL(k_scan_tiles_batch, dim3(gt), dim3(DEM_BLOCK), 0, st, rec,
                     (const int2*)b->d_map[MAP_CTILE], 0);
  hipLaunchKernelGGL(k_scan_apply_batch, dim3(gt), dim3(DEM_BLOCK), 0, st, rec,
                     (const int2*)b->d_map[MAP_CTILE], 0);
  hipLaunchKernelGGL(k_bin_scatter_batch, dim3(gp), dim3(DEM_BLOCK), 0, st, rec, mp);
  hipLaunchKernelGGL(k_cell_sort_batch, dim3(gc), dim3(DEM_BLOCK), 0, st, rec,
                     (const int2*)b->d_map[MAP_CELL]);
  hipLaunchKernelGGL(k_nbr_count_batch, dim3(gp), dim3(DEM_BLOCK), 0, st, rec, mp);
  hipLaunchKernelGGL(k_scan_tiles_batch, dim3(gs), dim3(DEM_BLOCK), 0, st, rec,
                     (const int2*)b->d_map[MAP_STILE], 1);
  hipLaunchKernelGGL(k_scan_apply_batch, dim3(gs), dim3(DEM_BLOCK), 0, st, rec,
                     (const int2*)b->d_map[MAP_STILE], 1);
  hipLaunchKernelGGL(k_nbr_fill_batch, dim3(gp), dim3(DEM_BLOCK), 0, st, rec, mp);
  hipLaunchKernelGGL(k_nbr_flip_batch, dim3(ns), dim3(DEM_MERGE_THREADS), 0, st, rec);
}

// ---- persistent stepping -----------------------------------------------------

template <bool MIXED, int FRONT>
static int persistent_occupancy(dem_batch_t b, int* blocks) {
  int per_cu = 0, dev = 0, cus = 0;
  BHIPCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                 &per_cu, reinterpret_cast<const void*>(&k_batch_persistent<MIXED, FRONT>),
                 DEM_BLOCK, 0));
  BHIPCHK(b, hipGetDevice(&dev));
  BHIPCHK(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  *blocks = per_cu * cus;
  return DEM_OK;
}

// Groups of the persistent launch: every system gets blocks in proportion to
// its pass blocks (at least one, at most one per pass block), within what
// the GPU holds resident at once for this very kernel.
static int persistent_groups(dem_batch_t b, int ns, bool mixed, int front) {
  std::vector<int> nb(ns);
  long long total = 0;
  for (int i = 0; i < ns; ++i) {
    nb[i] = nblocks(b->sys[i]->np, DEM_BLOCK);
    total += nb[i];
  }
  if (nb == b->pers_key && b->pers_grid > 0) return DEM_OK;
  int resident = 0;
  int rc = mixed ? (front ? persistent_occupancy<true, 1>(b, &resident)
                          : persistent_occupancy<true, 0>(b, &resident))
                 : (front ? persistent_occupancy<false, 1>(b, &resident)
                          : persistent_occupancy<false, 0>(b, &resident));
  if (rc) return rc;
  if (resident < ns)
    return bfail(b, DEM_ERR_CAPACITY,
                 "persistent stepping needs a resident block per system: %d systems, %d blocks",
                 ns, resident);
  const long long budget = std::min<long long>(resident, total);
  std::vector<int> g(ns);
  long long used = 0;
  for (int i = 0; i < ns; ++i) {
    g[i] = (int)std::max<long long>(1, std::min<long long>(nb[i], budget * nb[i] / total));
    used += g[i];
  }
  // what is left goes to the systems with the most pass blocks per block
  while (used < budget) {
    int best = -1;
    double worst = 0.0;
    for (int i = 0; i < ns; ++i) {
      if (g[i] >= nb[i]) continue;
      const double load = (double)nb[i] / g[i];
      if (load > worst) { worst = load; best = i; }
    }
    if (best < 0) break;
    ++g[best];
    ++used;
  }
  std::vector<int2> group;
  group.reserve((size_t)used);
  for (int i = 0; i < ns; ++i)
    for (int r = 0; r < g[i]; ++r) group.push_back(make_int2(i, r));
  if (group.size() > b->group_cap) {
    if (b->d_group) (void)hipFree(b->d_group);
    b->d_group = nullptr;
    BHIPCHK(b, hipMalloc((void**)&b->d_group, group.size() * sizeof(int2)));
    b->group_cap = group.size();
  }
  if ((size_t)ns > b->pers_cap) {
    if (b->d_gsize) (void)hipFree(b->d_gsize);
    if (b->d_gsync) (void)hipFree(b->d_gsync);
    if (b->d_scratch) (void)hipFree(b->d_scratch);
    if (b->d_scratch_mem) (void)hipFree(b->d_scratch_mem);
    b->d_gsize = nullptr;
    b->d_gsync = nullptr;
    b->d_scratch = nullptr;
    b->d_scratch_mem = nullptr;
    b->pers_cap = 0;
    const size_t per = (size_t)DEM_STALE_ROWS *
                       (sizeof(unsigned long long) * 2 + sizeof(unsigned short) + sizeof(int2) +
                        3 * sizeof(double2));
    BHIPCHK(b, hipMalloc((void**)&b->d_gsize, (size_t)ns * sizeof(int)));
    BHIPCHK(b, hipMalloc((void**)&b->d_gsync, (size_t)ns * sizeof(GroupSync)));
    BHIPCHK(b, hipMalloc((void**)&b->d_scratch, (size_t)ns * sizeof(MergeScratch)));
    BHIPCHK(b, hipMalloc((void**)&b->d_scratch_mem, (size_t)ns * per));
    std::vector<MergeScratch> ms((size_t)ns);
    for (int i = 0; i < ns; ++i) {
      char* p = b->d_scratch_mem + (size_t)i * per;
      const size_t R = DEM_STALE_ROWS;
      ms[i].sk = reinterpret_cast<unsigned long long*>(p);
      ms[i].key = reinterpret_cast<unsigned long long*>(p + R * 8);
      ms[i].hist = reinterpret_cast<double2*>(p + R * 16);
      ms[i].slot = reinterpret_cast<int2*>(p + R * 16 + R * 48);
      ms[i].si = reinterpret_cast<unsigned short*>(p + R * 16 + R * 48 + R * 8);
    }
    BHIPCHK(b, hipMemcpyAsync(b->d_scratch, ms.data(), ms.size() * sizeof(MergeScratch),
                              hipMemcpyHostToDevice, b->stream));
    b->pers_cap = (size_t)ns;
  }
  if (!b->d_abort) BHIPCHK(b, hipMalloc((void**)&b->d_abort, sizeof(int)));
  BHIPCHK(b, hipMemcpyAsync(b->d_group, group.data(), group.size() * sizeof(int2),
                            hipMemcpyHostToDevice, b->stream));
  BHIPCHK(b, hipMemcpyAsync(b->d_gsize, g.data(), (size_t)ns * sizeof(int),
                            hipMemcpyHostToDevice, b->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  b->pers_grid = (int)group.size();
  b->pers_key = nb;
  return DEM_OK;
}

static int batch_persistent_front(const dem_batch* b) {
  int blk[DEM_NSHAPES], wv[DEM_NSHAPES], fr[DEM_NSHAPES];
  shape_table(blk, wv, fr);
  // the persistent kernel walks the lists without LDS staging (bit 1)
  return (blk[b->shape] == DEM_BLOCK && wv[b->shape] == DEM_PASS_WAVES ? fr[b->shape] : fr[0]) & 1;
}

// nsteps steps of every system by persistent launches of persist_steps
// steps each.
static int persistent_steps(dem_batch_t b, int ns, bool mixed, int front, int64_t nsteps) {
  int rc = persistent_groups(b, ns, mixed, front);
  if (rc) return rc;
  hipStream_t st = b->stream;
  for (int64_t s = 0; s < nsteps;) {
    const int chunk = (int)std::min<int64_t>(b->persist_steps, nsteps - s);
    BHIPCHK(b, hipMemsetAsync(b->d_gsync, 0, (size_t)ns * sizeof(GroupSync), st));
    BHIPCHK(b, hipMemsetAsync(b->d_abort, 0, sizeof(int), st));
    PersistArgs pa;
    pa.group = b->d_group;
    pa.gsize = b->d_gsize;
    pa.gsync = b->d_gsync;
    pa.scratch = b->d_scratch;
    pa.abort = b->d_abort;
    pa.spin_limit = b->spin_limit;
    pa.nsteps = chunk;
    const BatchRec* rec = b->d_rec;
    void* args[] = {(void*)&rec, (void*)&pa};
    const void* fn = mixed ? (front ? reinterpret_cast<const void*>(&k_batch_persistent<true, 1>)
                                    : reinterpret_cast<const void*>(&k_batch_persistent<true, 0>))
                           : (front ? reinterpret_cast<const void*>(&k_batch_persistent<false, 1>)
                                    : reinterpret_cast<const void*>(&k_batch_persistent<false, 0>));
    BHIPCHK(b, hipLaunchCooperativeKernel(fn, dim3(b->pers_grid), dim3(DEM_BLOCK), args, 0, st));
    hipLaunchKernelGGL(k_persist_advance, dim3(nblocks(ns, 64)), dim3(64), 0, st, b->d_rec, ns,
                       chunk);
    int abort = 0;
    BHIPCHK(b, hipMemcpyAsync(&abort, b->d_abort, sizeof(int), hipMemcpyDeviceToHost, st));
    BHIPCHK(b, hipStreamSynchronize(st));
    if (abort)
      return bfail(b, DEM_ERR_DEVICE,
                   "persistent batch step: a group barrier wait ran out (the systems' states "
                   "are no longer valid)");
    s += chunk;
  }
  return DEM_OK;
}

static void drop_graphs(dem_batch_t b) {
  for (int k = 0; k < 2; ++k)
    if (b->gexec[k]) (void)hipGraphExecDestroy(b->gexec[k]);
  b->gexec[0] = b->gexec[1] = nullptr;
  b->gkey.clear();
}

// Captures nsteps batched steps from b->stream into an executable graph.
// The stream is captured in relaxed mode, so allocations other threads make
// meanwhile (systems being built for admission) are not disturbed.
static int capture_steps(dem_batch_t b, int nsteps, int ns, bool mixed, bool moving,
                         hipGraphExec_t* out) {
  hipStream_t st = b->stream;
  BHIPCHK(b, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  for (int k = 0; k < nsteps; ++k) enqueue_batch_step(b, ns, mixed, moving);
  hipGraph_t g = nullptr;
  const hipError_t le = hipGetLastError();
  const hipError_t ce = hipStreamEndCapture(st, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return bfail(b, DEM_ERR_DEVICE, "capturing the batched step failed: %s",
                 hipGetErrorString(le != hipSuccess ? le : ce));
  }
  const hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess)
    return bfail(b, DEM_ERR_DEVICE, "instantiating the batched step graph failed: %s",
                 hipGetErrorString(ie));
  return DEM_OK;
}

// The graphs of the current batch, captured again when what their launches
// depend on changed (the table and the maps, their sizes, the precision and
// the moving walls).
static int batch_graphs(dem_batch_t b, int ns, bool mixed, bool moving) {
  std::vector<long long> key = {(long long)(uintptr_t)b->d_rec, ns, mixed, moving,
                                b->graph_steps, b->shape, b->block_map};
  for (int k = 0; k < MAP_KINDS; ++k) {
    key.push_back((long long)(uintptr_t)b->d_map[k]);
    key.push_back(b->map_len[k]);
  }
  if (b->gexec[0] && key == b->gkey) return DEM_OK;
  drop_graphs(b);
  int rc = capture_steps(b, 1, ns, mixed, moving, &b->gexec[0]);
  if (rc == DEM_OK && b->graph_steps > 1)
    rc = capture_steps(b, b->graph_steps, ns, mixed, moving, &b->gexec[1]);
  if (rc) {
    drop_graphs(b);
    return rc;
  }
  b->gkey = key;
  return DEM_OK;
}

extern "C" {

int dem_batch_create(dem_batch_t* out) {
  if (!out) return bfail(nullptr, DEM_ERR_ARG, "null argument");
  *out = nullptr;
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1)
    return bfail(nullptr, DEM_ERR_DEVICE, "no HIP device available");
  dem_batch_t b = new dem_batch();
  b->err[0] = 0;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  for (int k = 0; k < dem_batch::K && e == hipSuccess; ++k) e = hipEventCreate(&b->ev[k]);
  if (e != hipSuccess) {
    dem_batch_destroy(b);
    return bfail(nullptr, DEM_ERR_DEVICE, "batch setup failed: %s", hipGetErrorString(e));
  }
  *out = b;
  return DEM_OK;
}

void dem_batch_destroy(dem_batch_t b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto h : b->sys) {
    h->in_batch = false;
    h->batch = nullptr;
  }
  for (int k = 0; k < dem_batch::K; ++k) {
    if (b->h_rec[k]) (void)hipHostFree(b->h_rec[k]);
    if (b->ev[k]) (void)hipEventDestroy(b->ev[k]);
  }
  for (int k = 0; k < MAP_KINDS; ++k)
    if (b->d_map[k]) (void)hipFree(b->d_map[k]);
  if (b->d_rec) (void)hipFree(b->d_rec);
  if (b->d_ctl) (void)hipFree(b->d_ctl);
  if (b->h_ctl) (void)hipHostFree(b->h_ctl);
  for (auto e : b->tev) (void)hipEventDestroy(e);
  if (b->d_group) (void)hipFree(b->d_group);
  if (b->d_gsize) (void)hipFree(b->d_gsize);
  if (b->d_gsync) (void)hipFree(b->d_gsync);
  if (b->d_scratch) (void)hipFree(b->d_scratch);
  if (b->d_scratch_mem) (void)hipFree(b->d_scratch_mem);
  if (b->d_abort) (void)hipFree(b->d_abort);
  drop_graphs(b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* dem_batch_last_error(dem_batch_t b) { return b ? b->err : g_err; }

int dem_batch_add(dem_batch_t b, dem_handle_t h) {
  if (!b || !h) return bfail(b, DEM_ERR_ARG, "null argument");
  if (!h->has_state || h->np == 0) return bfail(b, DEM_ERR_STATE, "upload a non-empty state first");
  if (h->in_batch) return bfail(b, DEM_ERR_STATE, "the system already belongs to a batch");
  if (h->fluid) return bfail(b, DEM_ERR_UNSUPPORTED, "systems with fluid flow are stepped alone");
  const int prec = h->mixed ? DEM_MIXED : DEM_FP64;
  if (b->precision >= 0 && prec != b->precision)
    return bfail(b, DEM_ERR_ARG, "all systems of a batch use the same precision mode");
  if (!h->rc_ready) {
    enqueue_refcell(h, h->cur);
    h->rc_ready = true;
  }
  BHIPCHK(b, hipStreamSynchronize(h->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  b->precision = prec;
  h->in_batch = true;
  h->batch = b;
  h->batch_first = true;
  b->sys.push_back(h);
  b->dirty = true;
  return DEM_OK;
}

int dem_batch_remove(dem_batch_t b, dem_handle_t h) {
  if (!b || !h) return bfail(b, DEM_ERR_ARG, "null argument");
  auto it = std::find(b->sys.begin(), b->sys.end(), h);
  if (it == b->sys.end()) return bfail(b, DEM_ERR_ARG, "the system is not in this batch");
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  b->sys.erase(it);
  h->in_batch = false;
  h->batch = nullptr;
  b->dirty = true;
  if (b->sys.empty()) b->precision = -1;
  return DEM_OK;
}

int dem_batch_size(dem_batch_t b) { return b ? (int)b->sys.size() : 0; }

int dem_batch_set_option(dem_batch_t b, const char* name, double value) {
  if (!b || !name) return bfail(b, DEM_ERR_ARG, "null argument");
  if (!strcmp(name, "timing")) {
    b->timing = value != 0.0;
  } else if (!strcmp(name, "graph")) {
    b->use_graph = value != 0.0;
  } else if (!strcmp(name, "pass_shape")) {
    if (!(value >= 0.0 && value < (double)DEM_NSHAPES) || value != (double)(int)value)
      return bfail(b, DEM_ERR_ARG, "pass_shape must be an integer in [0, %d)", DEM_NSHAPES);
    if ((int)value != b->shape) b->dirty = true;  // the pass-block map changes
    b->shape = (int)value;
  } else if (!strcmp(name, "block_map")) {
    if (value != 0.0 && value != 1.0) return bfail(b, DEM_ERR_ARG, "block_map must be 0 or 1");
    b->block_map = (int)value;
  } else if (!strcmp(name, "persistent")) {
    b->persistent = value != 0.0;
  } else if (!strcmp(name, "persist_steps")) {
    if (!(value >= 1.0 && value <= 100000.0))
      return bfail(b, DEM_ERR_ARG, "persist_steps must lie in [1, 100000]");
    b->persist_steps = (int)value;
  } else if (!strcmp(name, "spin_limit")) {
    if (!(value >= 0.0 && value <= 4.0e9))
      return bfail(b, DEM_ERR_ARG, "spin_limit must lie in [0, 4e9]");
    b->spin_limit = (unsigned)value;
  } else if (!strcmp(name, "graph_steps")) {
    if (!(value >= 1.0 && value <= 256.0))
      return bfail(b, DEM_ERR_ARG, "graph_steps must lie in [1, 256]");
    b->graph_steps = (int)value;
  } else {
    return bfail(b, DEM_ERR_ARG, "unknown option '%s'", name);
  }
  return DEM_OK;
}

int dem_batch_step(dem_batch_t b, int64_t nsteps, int64_t* steps_done, int32_t* escaped) {
  if (!b) return bfail(nullptr, DEM_ERR_ARG, "null batch");
  const size_t n = b->sys.size();
  for (size_t i = 0; i < n; ++i) {
    if (steps_done) steps_done[i] = 0;
    if (escaped) escaped[i] = 0;
  }
  if (n == 0 || nsteps <= 0) return DEM_OK;
  int rc = batch_capacity(b, n);
  if (rc) return rc;
  if (b->dirty) {
    rc = batch_maps(b);
    if (rc) return rc;
  }
  hipStream_t st = b->stream;
  const bool mixed = b->precision == DEM_MIXED;
  if (b->timing) {
    while (b->tev.size() < (size_t)nsteps + 1) {
      hipEvent_t e;
      BHIPCHK(b, hipEventCreate(&e));
      b->tev.push_back(e);
    }
  }
  const int ns = (int)n;
  // the table of this call, once; every step then advances it on the device
  const bool pers = batch_persistent_now(b);
  for (size_t i = 0; i < n; ++i) {
    rc = sync_halo(b->sys[i], !pers && shape_staged(b->shape), st);
    if (rc) return bfail(b, rc, "system %zu: %s", i, dem_last_error(b->sys[i]));
  }
  BatchRec* R = b->h_rec[0];
  if (b->ev_used[0]) BHIPCHK(b, hipEventSynchronize(b->ev[0]));
  for (size_t i = 0; i < n; ++i) batch_rec_init(b->sys[i], R[i]);
  BHIPCHK(b, hipMemcpyAsync(b->d_rec, R, n * sizeof(BatchRec), hipMemcpyHostToDevice, st));
  BHIPCHK(b, hipEventRecord(b->ev[0], st));
  b->ev_used[0] = true;
  BatchRec* rec = b->d_rec;
  bool moving = false;
  for (size_t i = 0; i < n; ++i) moving = moving || b->sys[i]->nw_move > 0;
  hipLaunchKernelGGL(k_reset_escape, dim3(nblocks(ns, 64)), dim3(64), 0, st, (const BatchRec*)rec,
                     ns);
  // the launch-per-phase path prepares each step's arguments in the records;
  // the persistent launch derives them from the records as uploaded
  if (!pers) hipLaunchKernelGGL(k_batch_prepare, dim3(nblocks(ns, 64)), dim3(64), 0, st, rec, ns);
  if (pers) {
    rc = persistent_steps(b, ns, mixed, batch_persistent_front(b), nsteps);
    if (rc) return rc;
  } else if (b->use_graph) {
    rc = batch_graphs(b, ns, mixed, moving);
    if (rc) return rc;
  }
  for (int64_t s = pers ? nsteps : 0; s < nsteps;) {
    if (b->use_graph && !b->timing && b->graph_steps > 1 && nsteps - s >= b->graph_steps) {
      BHIPCHK(b, hipGraphLaunch(b->gexec[1], st));
      s += b->graph_steps;
      continue;
    }
    if (b->timing) BHIPCHK(b, hipEventRecord(b->tev[s], st));
    if (b->use_graph) BHIPCHK(b, hipGraphLaunch(b->gexec[0], st));
    else enqueue_batch_step(b, ns, mixed, moving);
    s += 1;
  }
  if (b->timing) BHIPCHK(b, hipEventRecord(b->tev[nsteps], st));
  BHIPCHK(b, hipGetLastError());
  hipLaunchKernelGGL(k_gather_ctl, dim3(nblocks(ns, 64)), dim3(64), 0, st,
                     (const BatchRec*)b->d_rec, ns, b->d_ctl);
  BHIPCHK(b, hipMemcpyAsync(b->h_ctl, b->d_ctl, n * sizeof(Ctl), hipMemcpyDeviceToHost, st));
  BHIPCHK(b, hipStreamSynchronize(st));
  if (b->timing) {
    b->step_ms.assign((size_t)nsteps, 0.0);
    float ms;
    for (int64_t s = 0; s < nsteps; ++s)
      if (hipEventElapsedTime(&ms, b->tev[s], b->tev[s + 1]) == hipSuccess) b->step_ms[s] = ms;
  }
  int result = DEM_OK;
  for (size_t i = 0; i < n; ++i) {
    dem_handle_t h = b->sys[i];
    batch_advance_host(h, nsteps);
    const Ctl& c = b->h_ctl[i];
    int64_t done = nsteps;
    if (c.escaped > 0) {
      done = (int64_t)c.stop_step + 1;
      rollback_steps(h, done, nsteps);
      if (result == DEM_OK) {
        result = DEM_ERR_ESCAPE;
        bfail(b, DEM_ERR_ESCAPE, "system %zu: %d particles left the box in step %lld", i,
              c.escaped, (long long)done);
      }
    }
    h->steps += done;
    if (done > 0) h->time = (double)(h->next_step + done - 1) * h->dt;
    h->next_step += done;
    if (steps_done) steps_done[i] = done;
    if (escaped) escaped[i] = c.escaped;
    if (c.error) {
      const int erc = check_errors(h);
      if (erc && (result == DEM_OK || result == DEM_ERR_ESCAPE)) {
        result = erc;
        bfail(b, erc, "system %zu: %s", i, h->err);
      }
    }
  }
  return result;
}

int dem_batch_autotune(dem_batch_t b, int64_t steps_per_shape, dem_tune_t* out) {
  if (!b || !out) return bfail(b, DEM_ERR_ARG, "null argument");
  memset(out, 0, sizeof *out);
  if (steps_per_shape < 1) return bfail(b, DEM_ERR_ARG, "steps_per_shape must be at least 1");
  const size_t n = b->sys.size();
  if (n == 0) return bfail(b, DEM_ERR_STATE, "the batch has no systems");
  const bool timing = b->timing;
  const int keep = b->shape;
  const int map0 = b->block_map;
  std::vector<int64_t> done(n);
  // one segment, as in dem_autotune; *ms = the summed device time of the
  // timed steps (the whole batched step)
  auto seg = [&](int c, int m, int64_t nsteps, bool warm, double* ms) -> int {
    if (c != b->shape) b->dirty = true;
    b->shape = c;
    b->block_map = m;
    *ms = 0.0;
    if (warm) {
      b->timing = false;
      const int rc = dem_batch_step(b, 1, done.data(), nullptr);
      out->steps += done[0];
      if (rc != DEM_OK) return rc;
    }
    if (nsteps <= 0) return DEM_OK;
    b->timing = true;
    const int rc = dem_batch_step(b, nsteps, done.data(), nullptr);
    out->steps += done[0];
    if (rc == DEM_OK)
      for (double t : b->step_ms) *ms += t;
    return rc;
  };
  auto undo = [&](int rc) {
    b->timing = timing;
    if (b->shape != keep) b->dirty = true;
    b->shape = keep;
    b->block_map = map0;
    return rc;
  };
  const int64_t h1 = (steps_per_shape + 1) / 2, h2 = steps_per_shape - h1;
  double t[DEM_NSHAPES], x;
  for (int c = 0; c < DEM_NSHAPES; ++c) {
    const int rc = seg(c, map0, h1, true, &t[c]);
    if (rc != DEM_OK) return undo(rc);
  }
  for (int c = DEM_NSHAPES - 1; c >= 0; --c) {
    const int rc = seg(c, map0, h2, false, &x);
    if (rc != DEM_OK) return undo(rc);
    t[c] += x;
  }
  int best = 0;
  for (int c = 0; c < DEM_NSHAPES; ++c) {
    out->ms[c] = t[c] / (double)steps_per_shape;
    if (out->ms[c] < out->ms[best]) best = c;
  }
  shape_table(out->block, out->waves, out->front);
  // the two block orders in the best shape, timed in the order A B B A
  double tm[2] = {0.0, 0.0};
  const struct {
    int map;
    int64_t nsteps;
    bool warm;
  } abba[4] = {{map0, h1, false}, {1 - map0, h1, true}, {1 - map0, h2, false}, {map0, h2, false}};
  for (const auto& q : abba) {
    const int rc = seg(best, q.map, q.nsteps, q.warm, &x);
    if (rc != DEM_OK) return undo(rc);
    tm[q.map] += x;
  }
  out->ms_map[0] = tm[0] / (double)steps_per_shape;
  out->ms_map[1] = tm[1] / (double)steps_per_shape;
  if (best != b->shape) b->dirty = true;
  b->shape = best;
  b->block_map = (out->ms_map[1 - map0] < out->ms_map[map0]) ? 1 - map0 : map0;
  b->timing = timing;
  out->nshapes = DEM_NSHAPES;
  out->chosen = best;
  out->block_map = b->block_map;
  return DEM_OK;
}

// Device time of each step of the last dem_batch_step (option "timing").
int64_t dem_batch_step_times(dem_batch_t b, int64_t cap, double* ms) {
  if (!b) return 0;
  const int64_t n = (int64_t)b->step_ms.size();
  for (int64_t k = 0; k < n && k < cap && ms; ++k) ms[k] = b->step_ms[k];
  return n;
}

}  // extern "C"
