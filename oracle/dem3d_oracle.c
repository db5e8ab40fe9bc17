/*
 * dem3d_oracle.c -- CPU restatement of the DEM3D per-timestep path.
 *
 * TEST INFRASTRUCTURE ONLY (see dem3d_oracle.h).  Compiled with
 * -O2 -ffp-contract=off so that every multiply and add rounds on its own, as
 * the reference does when built without FMA contraction.
 *
 * Restated routines (reference file:line):
 *   InitCellStructure3D      src/MCells.f90:155-227
 *   FullNeighbourCells3D     src/MCells.f90:276-327
 *   AddAllNeighbours3D       src/MCells.f90:424-476  (+ VerifyPBC/ApplyPBC :1030-1052)
 *   CurrentCell3D            src/MCells.f90:498-513
 *   DistributeParticles3D    src/MCells.f90:542-564
 *   RedistributeParticles3D  src/MCells.f90:607-643  (TransferParticle :747-774)
 *   NeighbourParticles       src/MCells.f90:793-829
 *   Verify/Update/RemovePairContact  src/MContacts.f90:329-437
 *   Verify/Update/RemoveWallContact  src/MContacts.f90:441-549
 *   InitializeForce          src/MForces.f90:54-87 (tables arrive derived)
 *   ComputeAccelerations     src/MForces.f90:145-335
 *   CommonCalculations       src/MForces.f90:339-387
 *   ChangePlane              src/MForces.f90:391-406
 *   ConfigureWalls, Overlap  src/MConfig.f90:636-822
 *   InitializeAuxiliaries    src/MDEM.f90:110-162 (ABM coefficients)
 *   Euler, MoveParticlesEuler src/MDEM.f90:227-281
 *   MoveParticles            src/MDEM.f90:285-369
 *   ApplyPeriodicBoundaryConditions src/MDEM.f90:373-388
 *   UpdateParticleList (escape scan) src/MDEM.f90:612-626
 *   KineticEnergy            src/MDEM.f90:199-213
 *
 * Particles are 0-based here (reference id = index + 1); the relative order of
 * ids, which is all the reference's algorithms depend on, is unchanged.
 */
#include "dem3d_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXC DEM_MAX_COMPONENTS
#define MAXW DEM_MAX_WALLS

typedef dem_particle_t particle_t;

/* One node list of MContacts (TPartner chain), kept as an ascending array. */
typedef struct {
  int n, cap;
  int* idx;
  double* v1; /* 3 per entry: Vector1 (tangential force)    */
  double* v2; /* 3 per entry: Vector2 (tangential velocity) */
} plist_t;

typedef struct { double rap, sin_t, cos_t; } cone_t;
typedef struct { double max_alpha_u, min_alpha_w_sq, max_alpha_w_sq; } cone_comp_t;

struct orc_system {
  char err[512];
  /* constants */
  int nc, nw, pbc[3], order, nprev, n_euler;
  double lbox[3], invl[3], phalf[3], mhalf[3], g[3], dt;
  double radius[MAXC], mass[MAXC], inertia[MAXC], invm[MAXC], invi[MAXC],
      effm[MAXC];
  double mindij[MAXC * MAXC], mindijsq[MAXC * MAXC], kn[MAXC * MAXC],
      yn[MAXC * MAXC], kt[MAXC * MAXC], yt[MAXC * MAXC], musq[MAXC * MAXC];
  double knw[MAXC * MAXW], ynw[MAXC * MAXW], ktw[MAXC * MAXW],
      ytw[MAXC * MAXW], muwsq[MAXC * MAXW];
  double ab0, am0, ab[3], am[3];
  dem_wall_t wall[MAXW];
  int wall_order[MAXW];
  int walls_in_movement;           /* WallsInMovement (src/MDEM.f90:36) */
  /* fluid flow (mFluid) */
  int fluid, nl;
  double fluid_density, viscosity, flow_rate, area, delta_z, vf0, h_2, mh_2;
  double density[MAXC];
  double* vf;                      /* Vf(0:NL) */
  double* ef;                      /* Ef(0:NL) */
  dem_wall_motion_t wm[MAXW];      /* WM(1:NW) */
  int64_t next_step;
  cone_t cone[MAXW];
  cone_comp_t cone_comp[MAXW][MAXC];
  /* reference cell structure */
  int nx, ny, nz, nt, nxm1, nym1, nzm1;
  double nx_by_lx, ny_by_ly, nz_by_lz, nx_by2, ny_by2, nz_by2;
  int* cell_head;   /* [nt] first particle, -1 = null      */
  int* cell_next;   /* [np] next particle in the same cell */
  int* nbr_cells;   /* [nt*13] neighbour cells, list order */
  int* n_nbr_cells; /* [nt]                                */
  int* cell_index;  /* [np] 1-based CellIndex              */
  int cap;          /* particles the arrays hold (>= max_np) */
  int max_np;       /* MaxNP (src/MDEM.f90:22): size of MainParticle */
  /* particles and integrator */
  int np;
  particle_t* part;
  particle_t* pred;
  double* prev; /* Previous(j,i): [(j*np + i)*9 + c] */
  int ipc[3];
  double time;
  int has_state;
  /* contacts */
  plist_t* pairs;
  plist_t* walls;
  /* scratch */
  double* force;  /* [np*3] */
  double* torque; /* [np*3] */
  double* fscale; /* [np*2] */
  int* nbr_list;
  int nbr_cap;
  /* counters */
  int64_t ncol, evaluations, steps, pair_contacts, wall_contacts;
  double max_overlap;
};

static int fail(orc_handle_t h, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(h->err, sizeof h->err, fmt, ap);
  va_end(ap);
  return code;
}

/* ---- small vector helpers with the reference's association --------------- */

/* dot_product(a,b): accumulate from zero in index order (src/MDimension3D.f90:75) */
static inline double dot3(const double* a, const double* b) {
  double s = 0.0;
  s = s + a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}
/* sum(a*a) */
static inline double sumsq3(const double* a) { return dot3(a, a); }
/* CrossProduct (src/MDimension3D.f90:83-95) */
static inline void cross3(const double* a, const double* b, double* c) {
  double c0 = a[1] * b[2] - a[2] * b[1];
  double c1 = a[2] * b[0] - a[0] * b[2];
  double c2 = a[0] * b[1] - a[1] * b[0];
  c[0] = c0; c[1] = c1; c[2] = c2;
}
/* Fortran nint for real(8) -> default integer: round half away from zero. */
static inline double fnint(double x) { return (double)(int)lround(x); }

/* ---- contact lists (MContacts) -------------------------------------------- */

static void plist_reserve(plist_t* l, int need) {
  if (need <= l->cap) return;
  int cap = l->cap ? l->cap * 2 : 8;
  while (cap < need) cap *= 2;
  l->idx = (int*)realloc(l->idx, sizeof(int) * cap);
  l->v1 = (double*)realloc(l->v1, sizeof(double) * 3 * cap);
  l->v2 = (double*)realloc(l->v2, sizeof(double) * 3 * cap);
  l->cap = cap;
}
static void plist_free(plist_t* l) {
  free(l->idx); free(l->v1); free(l->v2);
  memset(l, 0, sizeof *l);
}
/* Position of the first entry with index >= key (the place where the
 * reference's walk in VerifyPairContact stops). */
static int plist_find(const plist_t* l, int key) {
  int k = 0;
  while (k < l->n && l->idx[k] < key) k++;
  return k;
}
static void plist_insert(plist_t* l, int pos, int key, const double* v1,
                         const double* v2) {
  plist_reserve(l, l->n + 1);
  memmove(l->idx + pos + 1, l->idx + pos, sizeof(int) * (l->n - pos));
  memmove(l->v1 + 3 * (pos + 1), l->v1 + 3 * pos, sizeof(double) * 3 * (l->n - pos));
  memmove(l->v2 + 3 * (pos + 1), l->v2 + 3 * pos, sizeof(double) * 3 * (l->n - pos));
  l->idx[pos] = key;
  memcpy(l->v1 + 3 * pos, v1, sizeof(double) * 3);
  memcpy(l->v2 + 3 * pos, v2, sizeof(double) * 3);
  l->n++;
}
static void plist_remove(plist_t* l, int pos) {
  memmove(l->idx + pos, l->idx + pos + 1, sizeof(int) * (l->n - pos - 1));
  memmove(l->v1 + 3 * pos, l->v1 + 3 * (pos + 1), sizeof(double) * 3 * (l->n - pos - 1));
  memmove(l->v2 + 3 * pos, l->v2 + 3 * (pos + 1), sizeof(double) * 3 * (l->n - pos - 1));
  l->n--;
}

/* State of one Verify*Contact call, consumed by Update/Remove
 * (the reference keeps it in module pointers CurrentPair/PreviousPair). */
typedef struct {
  plist_t* list;
  int pos;     /* position of the entry (if found) or of the insertion point */
  int key;
  int contacting;
} probe_t;

/* VerifyPairContact / VerifyWallContact (src/MContacts.f90:329-373,441-485) */
static probe_t verify(plist_t* list, int key, double* v1, double* v2) {
  probe_t p;
  p.list = list;
  p.key = key;
  p.pos = plist_find(list, key);
  p.contacting = p.pos < list->n && list->idx[p.pos] == key;
  if (p.contacting) {
    memcpy(v1, list->v1 + 3 * p.pos, sizeof(double) * 3);
    memcpy(v2, list->v2 + 3 * p.pos, sizeof(double) * 3);
  } else {
    v1[0] = v1[1] = v1[2] = 0.0;
    v2[0] = v2[1] = v2[2] = 0.0;
  }
  return p;
}
/* UpdatePairContact / UpdateWallContact (src/MContacts.f90:377-423,489-535) */
static void update(probe_t* p, const double* v1, const double* v2) {
  if (p->contacting) {
    memcpy(p->list->v1 + 3 * p->pos, v1, sizeof(double) * 3);
    memcpy(p->list->v2 + 3 * p->pos, v2, sizeof(double) * 3);
  } else {
    plist_insert(p->list, p->pos, p->key, v1, v2);
  }
}
/* RemovePairContact / RemoveWallContact (src/MContacts.f90:427-437,539-549) */
static void remove_entry(probe_t* p) { plist_remove(p->list, p->pos); }

/* ---- cell structure (MCells) ---------------------------------------------- */

/* VerifyPBC (src/MCells.f90:1030-1040) */
static void verify_pbc(int* ini, int* fin, int i, int m, int pbc) {
  *ini = -1;
  *fin = +1;
  if (!pbc) {
    if (i == 0) *ini = 0;
    if (i == m - 1) *fin = 0;
  }
}
/* ApplyPBC (src/MCells.f90:1044-1052) */
static int apply_pbc(int j, int m, int pbc) {
  if (pbc) {
    if (j == m) j = 0;
    if (j == -1) j = m - 1;
  }
  return j;
}

/* InitCellStructure3D (src/MCells.f90:155-227) and FullNeighbourCells3D
 * (src/MCells.f90:276-327, AddAllNeighbours3D :424-476, Method =
 * MolecularDynamics).  The neighbour list of each cell is head-pushed
 * (AddNeighbour :666-680), so it is stored in reverse insertion order. */
static int init_cells(orc_handle_t h, int max_ncell, const int* ref_grid) {
  double cutoff = 0.0;
  for (int c = 0; c < h->nc; c++)
    if (h->radius[c] > cutoff) cutoff = h->radius[c];
  cutoff = 2.0 * cutoff; /* InitializeCellList src/MForces.f90:96 */
  double lx = h->lbox[0], ly = h->lbox[1], lz = h->lbox[2];
  if (ref_grid && ref_grid[0] > 0 && ref_grid[1] > 0 && ref_grid[2] > 0) {
    h->nx = ref_grid[0]; h->ny = ref_grid[1]; h->nz = ref_grid[2];
  } else {
    int nx4 = (int)(lx / cutoff), ny4 = (int)(ly / cutoff), nz4 = (int)(lz / cutoff);
    long long n4 = (long long)nx4 * ny4 * nz4;
    if (n4 > max_ncell) {
      double lcell = pow(lx * ly * lz / (double)max_ncell, 1.0 / 3.0);
      h->nx = (int)(lx / lcell); h->ny = (int)(ly / lcell); h->nz = (int)(lz / lcell);
    } else {
      h->nx = nx4; h->ny = ny4; h->nz = nz4;
    }
  }
  if (h->nx < 1 || h->ny < 1 || h->nz < 1)
    return fail(h, DEM_ERR_ARG, "reference grid has an empty dimension (%d,%d,%d)",
                h->nx, h->ny, h->nz);
  h->nt = h->nx * h->ny * h->nz;
  h->nx_by_lx = (double)h->nx / lx;
  h->ny_by_ly = (double)h->ny / ly;
  h->nz_by_lz = (double)h->nz / lz;
  h->nx_by2 = (double)h->nx * 0.5;
  h->ny_by2 = (double)h->ny * 0.5;
  h->nz_by2 = (double)h->nz * 0.5;
  h->nxm1 = h->nx - 1; h->nym1 = h->ny - 1; h->nzm1 = h->nz - 1;
  h->cell_head = (int*)malloc(sizeof(int) * h->nt);
  h->nbr_cells = (int*)malloc(sizeof(int) * h->nt * 27);
  h->n_nbr_cells = (int*)calloc(h->nt, sizeof(int));
  for (int c = 0; c < h->nt; c++) h->cell_head[c] = -1;
  int nx = h->nx, ny = h->ny, nz = h->nz;
  for (int iz = 0; iz < nz; iz++) {
    int zi, zf;
    verify_pbc(&zi, &zf, iz, nz, h->pbc[2]);
    for (int iy = 0; iy < ny; iy++) {
      int yi, yf;
      verify_pbc(&yi, &yf, iy, ny, h->pbc[1]);
      for (int ix = 0; ix < nx; ix++) {
        int xi, xf;
        verify_pbc(&xi, &xf, ix, nx, h->pbc[0]);
        int icell = ix + nx * (iy + ny * iz);
        int tmp[27], nt = 0;
        for (int kz = zi; kz <= zf; kz++) {
          int jz = apply_pbc(iz + kz, nz, h->pbc[2]);
          for (int ky = yi; ky <= yf; ky++) {
            int jy = apply_pbc(iy + ky, ny, h->pbc[1]);
            for (int kx = xi; kx <= xf; kx++) {
              int kcell = kx + nx * (ky + ny * kz);
              if (kcell > 0) {
                int jx = apply_pbc(ix + kx, nx, h->pbc[0]);
                tmp[nt++] = jx + nx * (jy + ny * jz);
              }
            }
          }
        }
        for (int k = 0; k < nt; k++) h->nbr_cells[icell * 27 + k] = tmp[nt - 1 - k];
        h->n_nbr_cells[icell] = nt;
      }
    }
  }
  return DEM_OK;
}

/* CurrentCell3D (src/MCells.f90:498-513), 1-based result. */
static int current_cell(const orc_handle_t h, const double* r) {
  int ix = (int)(h->nx_by_lx * r[0] + h->nx_by2);
  int iy = (int)(h->ny_by_ly * r[1] + h->ny_by2);
  int iz = (int)(h->nz_by_lz * r[2] + h->nz_by2);
  ix = ix < h->nxm1 ? ix : h->nxm1; ix = ix > 0 ? ix : 0;
  iy = iy < h->nym1 ? iy : h->nym1; iy = iy > 0 ? iy : 0;
  iz = iz < h->nzm1 ? iz : h->nzm1; iz = iz > 0 ? iz : 0;
  return 1 + ix + (iy + iz * h->ny) * h->nx;
}

/* AddParticle (src/MCells.f90:684-698): push at head. */
static void add_particle(orc_handle_t h, int i, int icell1) {
  h->cell_next[i] = h->cell_head[icell1 - 1];
  h->cell_head[icell1 - 1] = i;
}

/* TransferParticle (src/MCells.f90:747-774). */
static void transfer_particle(orc_handle_t h, int i, int from1, int to1) {
  int* link = &h->cell_head[from1 - 1];
  while (*link != i) link = &h->cell_next[*link];
  *link = h->cell_next[i];
  add_particle(h, i, to1);
}

/* RedistributeParticles3D (src/MCells.f90:607-643). */
static void redistribute(orc_handle_t h, const particle_t* p) {
  for (int i = 0; i < h->np; i++) {
    int old = h->cell_index[i];
    int nw = current_cell(h, p[i].position);
    if (nw != old) {
      transfer_particle(h, i, old, nw);
      h->cell_index[i] = nw;
    }
  }
}

/* NeighbourParticles (src/MCells.f90:793-829), without the reference's fixed
 * 50-entry bound (src/MForces.f90:27): the list grows as needed. */
static int neighbour_particles(orc_handle_t h, int icell, int* npcc) {
  int n = 0;
  for (int q = h->cell_head[icell]; q >= 0; q = h->cell_next[q]) {
    if (n == h->nbr_cap) {
      h->nbr_cap *= 2;
      h->nbr_list = (int*)realloc(h->nbr_list, sizeof(int) * h->nbr_cap);
    }
    h->nbr_list[n++] = q;
  }
  *npcc = n;
  if (n == 0) return 0;
  for (int k = 0; k < h->n_nbr_cells[icell]; k++) {
    int jc = h->nbr_cells[icell * 27 + k];
    for (int q = h->cell_head[jc]; q >= 0; q = h->cell_next[q]) {
      if (n == h->nbr_cap) {
        h->nbr_cap *= 2;
        h->nbr_list = (int*)realloc(h->nbr_list, sizeof(int) * h->nbr_cap);
      }
      h->nbr_list[n++] = q;
    }
  }
  return n;
}

/* ---- walls (MConfig) ------------------------------------------------------- */

/* ConfigureWalls (src/MConfig.f90:636-688). */
static void configure_walls(orc_handle_t h) {
  int nws[3] = {0, 0, 0};
  for (int j = 0; j < h->nw; j++) {
    int s = h->wall[j].shape;
    nws[s] += 1;
    h->wall_order[j] = nws[s];
  }
  for (int j = 0; j < h->nw; j++) {
    if (h->wall[j].shape != DEM_SHAPE_CONE) continue;
    const dem_wall_t* w = &h->wall[j];
    double l_2 = w->l1, l = 2.0 * l_2, ra = w->l2, rp = w->l3;
    double aux = sqrt(l * l + (rp - ra) * (rp - ra));
    h->cone[j].rap = aux;
    h->cone[j].sin_t = l / aux;
    h->cone[j].cos_t = (rp - ra) / aux;
    for (int c = 0; c < h->nc; c++) {
      double ri = h->radius[c];
      double mn = (ra < rp ? ra : rp) - ri;
      mn = mn > 0.0 ? mn : 0.0;
      double mx = (ra > rp ? ra : rp) + ri;
      h->cone_comp[j][c].max_alpha_u = l_2 + ri;
      h->cone_comp[j][c].min_alpha_w_sq = mn * mn;
      h->cone_comp[j][c].max_alpha_w_sq = mx * mx;
    }
  }
}

/* Overlap (src/MConfig.f90:692-822).  Returns 1 on overlap, 0 otherwise, and
 * -1 for the degenerate cone case that needs the random generator
 * (Wij with AlphaW == 0, src/MConfig.f90:808-815), which is not restated. */
static int overlap(const orc_handle_t h, const particle_t* ip, int jw, double ri,
                   double* dij, double* nij, double* qji) {
  const dem_wall_t* w = &h->wall[jw];
  double rij[3];
  for (int k = 0; k < 3; k++) rij[k] = w->position[k] - ip->position[k];
  for (int k = 0; k < 3; k++)
    if (h->pbc[k]) rij[k] = rij[k] - h->lbox[k] * fnint(rij[k] * h->invl[k]);
  if (w->shape == DEM_SHAPE_RECTANGLE) {
    double alpha_n = dot3(rij, w->n);
    double abs_n = fabs(alpha_n);
    if (!(abs_n < ri)) return 0;
    double alpha_u = dot3(rij, w->u), alpha_w = dot3(rij, w->w);
    double abs_u = fabs(alpha_u), abs_w = fabs(alpha_w);
    if (!((abs_u < w->l1 + ri) && (abs_w < w->l2 + ri))) return 0;
    if ((abs_u < w->l1) && (abs_w < w->l2)) {
      double s = copysign(1.0, alpha_n);
      for (int k = 0; k < 3; k++) nij[k] = s * w->n[k];
      *dij = abs_n;
      for (int k = 0; k < 3; k++) qji[k] = -(alpha_u * w->u[k] + alpha_w * w->w[k]);
      return 1;
    }
    alpha_u = fmax(-w->l1, fmin(alpha_u, w->l1));
    alpha_w = fmax(-w->l2, fmin(alpha_w, w->l2));
    for (int k = 0; k < 3; k++) qji[k] = -(alpha_u * w->u[k] + alpha_w * w->w[k]);
    for (int k = 0; k < 3; k++) nij[k] = rij[k] + qji[k];
    double dsq = sumsq3(nij);
    if (!(dsq < ri * ri)) return 0;
    *dij = sqrt(dsq);
    for (int k = 0; k < 3; k++) nij[k] = nij[k] / *dij;
    return 1;
  }
  /* Cone */
  const cone_comp_t* cc = &h->cone_comp[jw][ip->index - 1];
  const cone_t* cn = &h->cone[jw];
  double alpha_u = dot3(rij, w->u);
  if (!(fabs(alpha_u) < cc->max_alpha_u)) return 0;
  double rijw[3];
  for (int k = 0; k < 3; k++) rijw[k] = rij[k] - alpha_u * w->u[k];
  double awsq = sumsq3(rijw);
  if (!((awsq >= cc->min_alpha_w_sq) && (awsq < cc->max_alpha_w_sq))) return 0;
  double alpha_w = sqrt(awsq);
  double lby2_au = w->l1 - alpha_u;
  double ra_aw = w->l2 - alpha_w;
  double lambda = lby2_au * cn->sin_t - ra_aw * cn->cos_t;
  double wv[3];
  if ((lambda >= 0.0) && (lambda <= cn->rap)) {
    double sd = -(lby2_au * cn->cos_t + ra_aw * cn->sin_t);
    *dij = fabs(sd);
    if (!(*dij < ri)) return 0;
    if (alpha_w == 0.0) return -1;
    for (int k = 0; k < 3; k++) wv[k] = rijw[k] / alpha_w;
    for (int k = 0; k < 3; k++) nij[k] = cn->sin_t * wv[k] + cn->cos_t * w->u[k];
    for (int k = 0; k < 3; k++) qji[k] = sd * nij[k] - rij[k];
    for (int k = 0; k < 3; k++) nij[k] = (qji[k] + rij[k]) / *dij;
    return 1;
  }
  lambda = fmax(0.0, fmin(lambda, cn->rap));
  if (alpha_w == 0.0) return -1;
  for (int k = 0; k < 3; k++) wv[k] = rijw[k] / alpha_w;
  double cu = cn->sin_t * lambda - lby2_au;
  double cw = cn->cos_t * lambda + ra_aw;
  for (int k = 0; k < 3; k++) nij[k] = cu * w->u[k] - cw * wv[k];
  double dsq = sumsq3(nij);
  if (!(dsq < ri * ri)) return 0;
  *dij = sqrt(dsq);
  for (int k = 0; k < 3; k++) qji[k] = nij[k] - rij[k];
  for (int k = 0; k < 3; k++) nij[k] = nij[k] / *dij;
  return 1;
}

/* ---- force evaluation (MForces) ------------------------------------------- */

/* ---- fluid drag (mFluid) ------------------------------------------------------ */

static const double PI_REF = 3.1415926535897932384626433832795; /* MUtil Pi */

/* UpdateVelocityField (src/mFluid.f90:96-140): porosity and fluid velocity
 * of the layers z_k = -H/2 + k*Delta_Z, k = 0..NL, from the evaluated
 * particles (their Index: Predicted%Index in the corrector pass). */
static void update_velocity_field(orc_handle_t h, const particle_t* p) {
  for (int k = 0; k <= h->nl; k++) {
    const double zk = h->mh_2 + (double)k * h->delta_z;
    double sa = 0.0, sb = 0.0, sc = 0.0;
    for (int i = 0; i < h->np; i++) {
      const int ind = p[i].index - 1;
      const double d = zk - p[i].position[2];
      const double diff2 = d * d;
      const double r2 = h->radius[ind] * h->radius[ind];
      if (diff2 < r2) {
        const double ai = PI_REF * (r2 - diff2);
        sa = sa + ai;
        const double dai = ai * h->density[ind];
        sb = sb + dai * p[i].velocity[2];
        sc = sc + dai;
      }
    }
    if (sa == 0.0) {
      h->vf[k] = h->vf0;
      h->ef[k] = 1.0;
    } else {
      const double es = sa / h->area;
      h->ef[k] = 1.0 - es;
      const double vs = sb / sc;
      h->vf[k] = (h->vf0 - es * vs) / h->ef[k];
    }
  }
}

/* DragForce (src/mFluid.f90:144-197): Di Felice's drag with the porosity
 * exponent found by fixed-point iteration.  Returns -1 when the particle is
 * outside the layers (the reference would read past Vf). */
static int drag_force(const orc_handle_t h, const particle_t* ip, double* fd) {
  const double zi = ip->position[2];
  const double jr = (zi + h->h_2) / h->delta_z;
  const int j = (int)jr;
  if (!(jr >= 0.0) || j + 1 > h->nl) return -1;
  double vfz = h->vf[j];
  vfz = vfz + (jr - (double)j) * (h->vf[j + 1] - vfz);
  double efz = h->ef[j];
  efz = efz + (jr - (double)j) * (h->ef[j + 1] - efz);
  double vr[3];
  for (int d = 0; d < 3; d++) vr[d] = -ip->velocity[d];
  vr[2] = vfz + vr[2];
  const double vra = sqrt(sumsq3(vr));
  if (vra > 0.0) {
    const double ri = h->radius[ip->index - 1];
    const double re = 2.0 * efz * ri * vra * h->fluid_density / h->viscosity;
    double n = 3.6, n0 = n + 1.0, x = 0.0;
    while (fabs(n - n0) > 1.e-8) {
      n0 = n;
      x = re / pow(efz, n);
      /* x**0.75 as the reference's compiler evaluates it: LLVM rewrites
       * pow(x, 0.75) into sqrt(x)*sqrt(sqrt(x)) (within an ulp of pow) */
      const double sx = sqrt(x);
      const double x075 = sx * sqrt(sx);
      n = (4.8 + 0.42 * x075) / (1.0 + 0.175 * x075);
    }
    double cd1 = 0.63 + 4.8 / sqrt(x);
    cd1 = cd1 * cd1;
    double cd = 0.63 + 4.8 / sqrt(re);
    cd = cd * cd;
    const double fi = (cd1 / cd) * pow(efz, 2.0 * (1.0 - n));
    const double c = cd * (PI_REF / 2.0) * ri * ri * h->fluid_density * vra;
    for (int d = 0; d < 3; d++) fd[d] = c * vr[d] * fi;
  } else {
    fd[0] = fd[1] = fd[2] = 0.0;
  }
  return 0;
}

/* ChangePlane (src/MForces.f90:391-406) */
static void change_plane(const double* v, const double* n, double* out) {
  double vn = dot3(v, n);
  double p[3];
  for (int k = 0; k < 3; k++) p[k] = v[k] - vn * n[k];
  double nsq = sumsq3(p);
  if (nsq != 0.0) {
    double s = sqrt(sumsq3(v) / nsq);
    for (int k = 0; k < 3; k++) out[k] = s * p[k];
  } else {
    out[0] = out[1] = out[2] = 0.0;
  }
}

/* Module-scope temporaries of MForces (src/MForces.f90:36-44) that
 * CommonCalculations reads and writes. */
typedef struct {
  double ri, nij[3], vij[3], vijc[3], disp_n[3], ft0[3], vt0[3], delta_dsq,
      half_dt;
  int contacting;
  /* outputs */
  double fn[3], ft[3], vt[3], n_x_ft[3];
  double i_force[3], i_torque[3];
} pair_ctx_t;

/* CommonCalculations (src/MForces.f90:339-387) */
static void common_calculations(pair_ctx_t* c, const double* rij, double kn,
                                double yn, double kt, double yt, double musq) {
  double vn_s = dot3(c->vijc, c->nij);
  double vijn[3];
  for (int k = 0; k < 3; k++) vijn[k] = vn_s * c->nij[k];
  for (int k = 0; k < 3; k++) c->vt[k] = c->vijc[k] - vijn[k];
  for (int k = 0; k < 3; k++) c->fn[k] = kn * c->disp_n[k] + yn * vijn[k];
  if (c->contacting) {
    double vt0p[3], ft0p[3], ddt[3], dvt[3];
    change_plane(c->vt0, c->nij, vt0p);
    for (int k = 0; k < 3; k++) c->vt0[k] = vt0p[k];
    for (int k = 0; k < 3; k++) ddt[k] = c->half_dt * (c->vt0[k] + c->vt[k]);
    for (int k = 0; k < 3; k++) dvt[k] = c->vt[k] - c->vt0[k];
    change_plane(c->ft0, c->nij, ft0p);
    for (int k = 0; k < 3; k++) c->ft[k] = ft0p[k] + kt * ddt[k] + yt * dvt[k];
  } else {
    double vsq = sumsq3(c->vij);
    double pdt;
    if (vsq > 0.0) {
      double rv = dot3(rij, c->vij);
      double t = (rv + sqrt(rv * rv + vsq * c->delta_dsq)) / vsq;
      double two_half = 2.0 * c->half_dt;
      pdt = t < two_half ? t : two_half; /* Fortran min(a,b) */
    } else {
      pdt = 0.0;
    }
    double f = kt * pdt + yt;
    for (int k = 0; k < 3; k++) c->ft[k] = f * c->vt[k];
  }
  double mu_fn_sq = musq * sumsq3(c->fn);
  double ft_sq = sumsq3(c->ft);
  if (ft_sq > mu_fn_sq) {
    double s = sqrt(mu_fn_sq / ft_sq);
    for (int k = 0; k < 3; k++) c->ft[k] = s * c->ft[k];
  }
  for (int k = 0; k < 3; k++) c->i_force[k] = c->i_force[k] + c->fn[k] + c->ft[k];
  cross3(c->nij, c->ft, c->n_x_ft);
  double rn[3], t[3];
  for (int k = 0; k < 3; k++) rn[k] = c->ri * c->nij[k];
  cross3(rn, c->ft, t);
  for (int k = 0; k < 3; k++) c->i_torque[k] = c->i_torque[k] + t[k];
}

/* ComputeAccelerations (src/MForces.f90:145-335) on particle array p. */
static int compute_accelerations(orc_handle_t h, particle_t* p, double time_step,
                                 int upd) {
  const int np = h->np;
  pair_ctx_t c;
  memset(&c, 0, sizeof c);
  c.half_dt = 0.5 * time_step;
  redistribute(h, p);
  memset(h->force, 0, sizeof(double) * 3 * np);
  memset(h->torque, 0, sizeof(double) * 3 * np);
  memset(h->fscale, 0, sizeof(double) * 2 * np);
  int64_t npairs = 0, nwallc = 0;
  double max_ov = 0.0;
  if (h->fluid) update_velocity_field(h, p);
  for (int icell = 0; icell < h->nt; icell++) {
    int npcc;
    int npsc = neighbour_particles(h, icell, &npcc);
    const int* lst = h->nbr_list;
    for (int ii = 0; ii < npcc; ii++) {
      const int i = lst[ii];
      const particle_t* ip = &p[i];
      const int ci = ip->index - 1;
      c.ri = h->radius[ci];
      for (int k = 0; k < 3; k++) c.i_force[k] = h->effm[ci] * h->g[k];
      c.i_torque[0] = c.i_torque[1] = c.i_torque[2] = 0.0;
      h->fscale[2 * i] += fabs(h->effm[ci]) * sqrt(sumsq3(h->g));
      for (int jj = ii + 1; jj < npsc; jj++) {
        const int j = lst[jj];
        const particle_t* jp = &p[j];
        const int cj = jp->index - 1;
        const double rj = h->radius[cj];
        double rij[3];
        for (int k = 0; k < 3; k++) rij[k] = jp->position[k] - ip->position[k];
        for (int k = 0; k < 3; k++)
          if (h->pbc[k]) rij[k] = rij[k] - h->lbox[k] * fnint(rij[k] * h->invl[k]);
        double dsq = sumsq3(rij);
        int imin = i < j ? i : j, imax = i < j ? j : i;
        probe_t pr = verify(&h->pairs[imin], imax, c.ft0, c.vt0);
        c.contacting = pr.contacting;
        c.delta_dsq = h->mindijsq[ci + cj * h->nc] - dsq;
        if (c.delta_dsq > 0.0) {
          if (!c.contacting) h->ncol++;
          npairs++;
          double dij = sqrt(dsq);
          for (int k = 0; k < 3; k++) c.nij[k] = rij[k] / dij;
          double md = h->mindij[ci + cj * h->nc];
          double ov = (md - dij) / md;
          if (ov > max_ov) max_ov = ov;
          for (int k = 0; k < 3; k++) c.disp_n[k] = (dij - md) * c.nij[k];
          for (int k = 0; k < 3; k++) c.vij[k] = jp->velocity[k] - ip->velocity[k];
          double ws[3], wx[3];
          for (int k = 0; k < 3; k++) ws[k] = c.ri * ip->omega[k] + rj * jp->omega[k];
          cross3(ws, c.nij, wx);
          for (int k = 0; k < 3; k++) c.vijc[k] = c.vij[k] - wx[k];
          const int t = ci + cj * h->nc;
          common_calculations(&c, rij, h->kn[t], h->yn[t], h->kt[t], h->yt[t],
                              h->musq[t]);
          double* fj = &h->force[3 * j];
          double* tj = &h->torque[3 * j];
          for (int k = 0; k < 3; k++) fj[k] = fj[k] - c.fn[k] - c.ft[k];
          for (int k = 0; k < 3; k++) tj[k] = tj[k] + rj * c.n_x_ft[k];
          double a = sqrt(sumsq3(c.fn)) + sqrt(sumsq3(c.ft));
          h->fscale[2 * i] += a;
          h->fscale[2 * j] += a;
          h->fscale[2 * i + 1] += c.ri * sqrt(sumsq3(c.ft));
          h->fscale[2 * j + 1] += rj * sqrt(sumsq3(c.ft));
          if (upd) update(&pr, c.ft, c.vt);
        } else {
          if (c.contacting && upd) remove_entry(&pr);
        }
      }
      for (int jw = 0; jw < h->nw; jw++) {
        const dem_wall_t* w = &h->wall[jw];
        probe_t pr = verify(&h->walls[i], jw, c.ft0, c.vt0);
        c.contacting = pr.contacting;
        double dij, qji[3];
        int ov = overlap(h, ip, jw, c.ri, &dij, c.nij, qji);
        if (ov < 0)
          return fail(h, DEM_ERR_UNSUPPORTED,
                      "particle %d lies on the axis of cone wall %d", i + 1, jw + 1);
        if (ov) {
          nwallc++;
          c.delta_dsq = c.ri * c.ri - dij * dij;
          for (int k = 0; k < 3; k++) c.disp_n[k] = (dij - c.ri) * c.nij[k];
          double wq[3];
          cross3(w->omega, qji, wq);
          for (int k = 0; k < 3; k++) c.vij[k] = w->velocity[k] + wq[k] - ip->velocity[k];
          double rw[3], wx[3];
          for (int k = 0; k < 3; k++) rw[k] = c.ri * ip->omega[k];
          cross3(rw, c.nij, wx);
          for (int k = 0; k < 3; k++) c.vijc[k] = c.vij[k] - wx[k];
          double rc[3];
          for (int k = 0; k < 3; k++) rc[k] = w->position[k] + qji[k] - ip->position[k];
          const int t = ci + jw * h->nc;
          common_calculations(&c, rc, h->knw[t], h->ynw[t], h->ktw[t], h->ytw[t],
                              h->muwsq[t]);
          h->fscale[2 * i] += sqrt(sumsq3(c.fn)) + sqrt(sumsq3(c.ft));
          h->fscale[2 * i + 1] += c.ri * sqrt(sumsq3(c.ft));
          if (upd) update(&pr, c.ft, c.vt);
        } else {
          if (c.contacting && upd) remove_entry(&pr);
        }
      }
      if (h->fluid) {
        double fd[3];
        if (drag_force(h, ip, fd))
          return fail(h, DEM_ERR_UNSUPPORTED, "particle %d is outside the fluid layers", i + 1);
        for (int k = 0; k < 3; k++) c.i_force[k] = c.i_force[k] + fd[k];
      }
      double* fi = &h->force[3 * i];
      double* ti = &h->torque[3 * i];
      for (int k = 0; k < 3; k++) fi[k] = fi[k] + c.i_force[k];
      for (int k = 0; k < 3; k++) ti[k] = ti[k] + c.i_torque[k];
    }
  }
  for (int i = 0; i < np; i++) {
    const int ci = p[i].index - 1;
    for (int k = 0; k < 3; k++) p[i].acceleration[k] = h->force[3 * i + k] * h->invm[ci];
    for (int k = 0; k < 3; k++) p[i].alpha[k] = h->torque[3 * i + k] * h->invi[ci];
  }
  h->pair_contacts = npairs;
  h->wall_contacts = nwallc;
  h->max_overlap = max_ov;
  h->evaluations++;
  return DEM_OK;
}

/* ---- integrator (MDEM) ------------------------------------------------------ */

/* ApplyPeriodicBoundaryConditions (src/MDEM.f90:373-388) */
static void apply_pbc_particles(orc_handle_t h, particle_t* p) {
  for (int d = 0; d < 3; d++) {
    if (!h->pbc[d]) continue;
    for (int i = 0; i < h->np; i++) {
      double pos = p[i].position[d];
      p[i].position[d] = pos - h->lbox[d] * fnint(pos * h->invl[d]);
    }
  }
}

/* ipc = cshift(ipc, shift=-1) */
static void rotate_ipc(orc_handle_t h) {
  if (h->nprev == 0) return;
  int last = h->ipc[h->nprev - 1];
  for (int k = h->nprev - 1; k > 0; k--) h->ipc[k] = h->ipc[k - 1];
  h->ipc[0] = last;
}

static double* prev_slot(orc_handle_t h, int slot1, int i) {
  return &h->prev[((size_t)i * 3 + (slot1 - 1)) * 9];
}

static void store_ring(orc_handle_t h) {
  for (int i = 0; i < h->np; i++) {
    double* s = prev_slot(h, h->ipc[0], i);
    memcpy(s, h->part[i].velocity, sizeof(double) * 3);
    memcpy(s + 3, h->part[i].acceleration, sizeof(double) * 3);
    memcpy(s + 6, h->part[i].alpha, sizeof(double) * 3);
  }
}

/* MoveWalls (src/MDEM.f90:392-408).  Phase - TwoPi*real(int(Phase/TwoPi))
 * reduces the phase with truncation toward zero. */
static void move_walls(orc_handle_t h, double time) {
  const double two_pi = 2.0 * 3.14159265358979323846;
  for (int j = 0; j < h->nw; j++) {
    const dem_wall_motion_t* m = &h->wm[j];
    for (int d = 0; d < 3; d++) {
      double phase = m->two_pi_frequency[d] * time + m->phase_angle[d];
      phase = phase - two_pi * (double)(int32_t)(phase / two_pi);
      h->wall[j].position[d] = m->constant_factor[d] + m->amplitude[d] * cos(phase);
      h->wall[j].velocity[d] = m->velocity_amplitude[d] * sin(phase);
    }
  }
}

/* MoveParticles (src/MDEM.f90:285-369) */
static int move_particles(orc_handle_t h, int64_t step) {
  int rc = compute_accelerations(h, h->part, h->dt, 1);
  if (rc) return rc;
  for (int i = 0; i < h->np; i++) {
    double sv[3] = {0, 0, 0}, sa[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
    for (int k = 0; k < h->nprev; k++) {
      const double abk = h->ab[k];
      const double* pv = prev_slot(h, h->ipc[k], i);
      for (int d = 0; d < 3; d++) sv[d] = sv[d] + abk * pv[d];
      for (int d = 0; d < 3; d++) sa[d] = sa[d] + abk * pv[3 + d];
      for (int d = 0; d < 3; d++) sl[d] = sl[d] + abk * pv[6 + d];
    }
    particle_t* q = &h->part[i];
    particle_t* r = &h->pred[i];
    /* Predicted(i)%Index is not assigned here: it keeps what InsertParticle,
     * ResizeVariables or ReadIntegrationVariables last set, which after a
     * removal can be the component of the particle that left slot i. */
    for (int d = 0; d < 3; d++) r->position[d] = q->position[d] + h->ab0 * q->velocity[d] + sv[d];
    for (int d = 0; d < 3; d++) r->velocity[d] = q->velocity[d] + h->ab0 * q->acceleration[d] + sa[d];
    for (int d = 0; d < 3; d++) r->omega[d] = q->omega[d] + h->ab0 * q->alpha[d] + sl[d];
  }
  apply_pbc_particles(h, h->pred);
  if (h->nprev > 0) {
    rotate_ipc(h);
    store_ring(h);
  }
  h->time = (double)step * h->dt;
  if (h->walls_in_movement) move_walls(h, h->time);
  rc = compute_accelerations(h, h->pred, h->dt, 0);
  if (rc) return rc;
  for (int i = 0; i < h->np; i++) {
    double sv[3] = {0, 0, 0}, sa[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
    for (int k = 0; k < h->nprev; k++) {
      const double amk = h->am[k];
      const double* pv = prev_slot(h, h->ipc[k], i);
      for (int d = 0; d < 3; d++) sv[d] = sv[d] + amk * pv[d];
      for (int d = 0; d < 3; d++) sa[d] = sa[d] + amk * pv[3 + d];
      for (int d = 0; d < 3; d++) sl[d] = sl[d] + amk * pv[6 + d];
    }
    particle_t* q = &h->part[i];
    const particle_t* r = &h->pred[i];
    for (int d = 0; d < 3; d++) q->position[d] = q->position[d] + h->am0 * r->velocity[d] + sv[d];
    for (int d = 0; d < 3; d++) q->velocity[d] = q->velocity[d] + h->am0 * r->acceleration[d] + sa[d];
    for (int d = 0; d < 3; d++) q->omega[d] = q->omega[d] + h->am0 * r->alpha[d] + sl[d];
  }
  apply_pbc_particles(h, h->part);
  h->steps++;
  return DEM_OK;
}

/* Euler and MoveParticlesEuler (src/MDEM.f90:227-281) */
static int move_particles_euler(orc_handle_t h) {
  if (h->nprev == 0) return fail(h, DEM_ERR_STATE, "Euler start-up needs MethodOrder > 1");
  double hstep = h->dt / (double)h->n_euler;
  rotate_ipc(h);
  store_ring(h);
  for (int s = 0; s < h->n_euler; s++) {
    double hsq2 = 0.5 * (hstep * hstep);
    int rc = compute_accelerations(h, h->part, hstep, 1);
    if (rc) return rc;
    for (int i = 0; i < h->np; i++) {
      particle_t* q = &h->part[i];
      for (int d = 0; d < 3; d++)
        q->position[d] = q->position[d] + q->velocity[d] * hstep + q->acceleration[d] * hsq2;
      for (int d = 0; d < 3; d++) q->velocity[d] = q->velocity[d] + q->acceleration[d] * hstep;
      for (int d = 0; d < 3; d++) q->omega[d] = q->omega[d] + q->alpha[d] * hstep;
    }
    h->time = h->time + hstep;
    if (h->walls_in_movement) move_walls(h, h->time);
  }
  apply_pbc_particles(h, h->part);
  return DEM_OK;
}

/* Escape test of UpdateParticleList (src/MDEM.f90:618-626) */
static int count_escaped(orc_handle_t h) {
  int n = 0;
  for (int i = 0; i < h->np; i++) {
    const double* r = h->part[i].position;
    int out = 0;
    for (int d = 0; d < 3; d++) out |= (r[d] > h->phalf[d]) || (r[d] < h->mhalf[d]);
    n += out;
  }
  return n;
}

/* ---- particle number changes (MDEM, MCells, MContacts) --------------------- */

/* Room for n particles in every per-particle array, contents kept. */
static void reserve_particles(orc_handle_t h, int n) {
  if (n <= h->cap) return;
  const int old = h->cap;
  const size_t c = (size_t)n;
  h->part = (particle_t*)realloc(h->part, sizeof(particle_t) * c);
  h->pred = (particle_t*)realloc(h->pred, sizeof(particle_t) * c);
  h->prev = (double*)realloc(h->prev, sizeof(double) * 27 * c);
  h->cell_next = (int*)realloc(h->cell_next, sizeof(int) * c);
  h->cell_index = (int*)realloc(h->cell_index, sizeof(int) * c);
  h->force = (double*)realloc(h->force, sizeof(double) * 3 * c);
  h->torque = (double*)realloc(h->torque, sizeof(double) * 3 * c);
  h->fscale = (double*)realloc(h->fscale, sizeof(double) * 2 * c);
  h->pairs = (plist_t*)realloc(h->pairs, sizeof(plist_t) * c);
  h->walls = (plist_t*)realloc(h->walls, sizeof(plist_t) * c);
  memset(h->pairs + old, 0, sizeof(plist_t) * (c - old));
  memset(h->walls + old, 0, sizeof(plist_t) * (c - old));
  memset(h->part + old, 0, sizeof(particle_t) * (c - old));
  memset(h->pred + old, 0, sizeof(particle_t) * (c - old));
  memset(h->prev + 27 * (size_t)old, 0, sizeof(double) * 27 * (c - old));
  h->cap = n;
}

/* ResizeVariables / ResizeContactList (src/MDEM.f90:566-608,
 * src/MContacts.f90:53-71): MaxNP becomes m and MainParticle keeps the lists
 * of particles 1..NP only (the new elements start null). */
static void resize_variables(orc_handle_t h, int m) {
  for (int i = h->np; i < h->cap; i++) {
    plist_free(&h->pairs[i]);
    plist_free(&h->walls[i]);
  }
  h->max_np = m;
  reserve_particles(h, m);
  for (int i = 0; i < h->np; i++) h->pred[i].index = h->part[i].index;
}

/* The Monte Carlo neighbour list of a cell (0-based): HinderNeighbourCellsList,
 * NeighbourCells(iCell, MonteCarlo, ...) and AllNeighbourParticles
 * (src/MCells.f90:340-374,424-476,833-860,1056-1074).  The cell's own list,
 * then the particles of each of the (up to 26) cells around it, every kCell /= 0
 * of AddAllNeighbours3D, in the head-pushed (reversed) order of AddNeighbour.
 * 0-based particle indices into out (grown as needed); returns the count. */
static int mc_neighbour_particles(orc_handle_t h, int icell, int** out, int* cap) {
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  const int iz = icell / (nx * ny);
  const int iy = icell / nx - iz * ny;
  const int ix = icell - (iy + iz * ny) * nx;
  int xi, xf, yi, yf, zi, zf;
  verify_pbc(&xi, &xf, ix, nx, h->pbc[0]);
  verify_pbc(&yi, &yf, iy, ny, h->pbc[1]);
  verify_pbc(&zi, &zf, iz, nz, h->pbc[2]);
  int cells[27], nc = 0;
  for (int kz = zi; kz <= zf; kz++) {
    const int jz = apply_pbc(iz + kz, nz, h->pbc[2]);
    for (int ky = yi; ky <= yf; ky++) {
      const int jy = apply_pbc(iy + ky, ny, h->pbc[1]);
      for (int kx = xi; kx <= xf; kx++) {
        if (kx + nx * (ky + ny * kz) == 0) continue;
        const int jx = apply_pbc(ix + kx, nx, h->pbc[0]);
        cells[nc++] = jx + nx * (jy + ny * jz);
      }
    }
  }
  int n = 0;
  for (int k = -1; k < nc; k++) {
    const int c = k < 0 ? icell : cells[nc - 1 - k];
    for (int q = h->cell_head[c]; q >= 0; q = h->cell_next[q]) {
      if (n == *cap) {
        *cap = *cap ? 2 * *cap : 256;
        *out = (int*)realloc(*out, sizeof(int) * *cap);
      }
      (*out)[n++] = q;
    }
  }
  return n;
}

/* RemoveParticleFromCell (src/MCells.f90:702-724). */
static void remove_from_cell(orc_handle_t h, int i, int icell) {
  int* link = &h->cell_head[icell];
  while (*link != i) link = &h->cell_next[*link];
  *link = h->cell_next[i];
}

/* ReplaceParticleIndexInCell (src/MCells.f90:728-743): the entry keeps its
 * place in the list. */
static void replace_in_cell(orc_handle_t h, int old, int icell, int now) {
  int* link = &h->cell_head[icell];
  while (*link != old) link = &h->cell_next[*link];
  *link = now;
  h->cell_next[now] = h->cell_next[old];
}

/* RemoveParticleFromContactList (src/MContacts.f90:579-633). */
static void remove_from_contacts(orc_handle_t h, int i, const int* nb, int n) {
  for (int jj = 0; jj < n; jj++) {
    plist_t* l = &h->pairs[nb[jj]];
    if (l->n == 0) continue;
    int pos = 0;
    while (l->idx[pos] < i && pos + 1 < l->n) pos++;
    if (l->idx[pos] == i) plist_remove(l, pos);
  }
  h->pairs[i].n = 0;
  h->walls[i].n = 0;
}

/* ChangeIndexInContactList (src/MContacts.f90:173-252): particle NP (here
 * old) becomes i.  In the lists of the neighbours j < i, a last entry old
 * is renamed i and moved behind the last earlier entry below i; a last entry
 * old in the list of a neighbour j > i moves to the list of i as partner j,
 * in neighbour order.  Wall contacts move with the particle. */
static void change_index_in_contacts(orc_handle_t h, int old, int i, const int* nb, int n) {
  plist_t first;
  memset(&first, 0, sizeof first);
  for (int jj = 0; jj < n; jj++) {
    const int j = nb[jj];
    plist_t* l = &h->pairs[j];
    if (l->n == 0) continue;
    const int last = l->n - 1;
    if (l->idx[last] != old) continue;
    double v1[3], v2[3];
    memcpy(v1, l->v1 + 3 * last, sizeof v1);
    memcpy(v2, l->v2 + 3 * last, sizeof v2);
    if (j < i) {
      if (last == 0) {
        l->idx[0] = i;
      } else {
        int point = -1;
        for (int k = 0; k < last; k++)
          if (l->idx[k] < i) point = k;
        plist_remove(l, last);
        plist_insert(l, point + 1, i, v1, v2);
      }
    } else {
      plist_remove(l, last);
      plist_insert(&first, first.n, j, v1, v2);
    }
  }
  plist_free(&h->pairs[i]);
  h->pairs[i] = first;
  plist_free(&h->walls[i]);
  h->walls[i] = h->walls[old];
  memset(&h->walls[old], 0, sizeof(plist_t));
}

/* UpdateParticleList (src/MDEM.f90:612-671): particles outside the box are
 * removed, i = NP down to 1; the last particle takes the place of each. */
int orc_remove_escaped(orc_handle_t h, int32_t cap, int32_t* removed, int32_t* nremoved) {
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  int count = 0;
  int* nb = NULL;
  int nbcap = 0;
  for (int i = h->np - 1; i >= 0; i--) {
    const double* r = h->part[i].position;
    int out = 0;
    for (int d = 0; d < 3; d++) out |= (r[d] > h->phalf[d]) || (r[d] < h->mhalf[d]);
    if (!out) continue;
    const int icell = h->cell_index[i] - 1;
    const int n = mc_neighbour_particles(h, icell, &nb, &nbcap);
    remove_from_cell(h, i, icell);
    remove_from_contacts(h, i, nb, n);
    const int last = h->np - 1;
    if (i != last) {
      h->part[i] = h->part[last];
      memcpy(&h->prev[(size_t)i * 27], &h->prev[(size_t)last * 27], sizeof(double) * 27);
      replace_in_cell(h, last, h->cell_index[last] - 1, i);
      h->cell_index[i] = h->cell_index[last];
      change_index_in_contacts(h, last, i, nb, n);
    }
    h->np -= 1;
    if (removed && count < cap) removed[count] = i + 1;
    count++;
    if (h->max_np >= (int)(1.1 * (double)h->np)) resize_variables(h, (int)(1.05 * (double)h->np));
  }
  free(nb);
  if (nremoved) *nremoved = count;
  return DEM_OK;
}

/* ranf (src/MUtil.f90:16-44): Park-Miller "minimal standard" generator by
 * Schrage's method combined with a Marsaglia shift sequence.  aam is
 * nearest(1.0,-1.0)/IM evaluated in default (single) precision. */
double orc_ranf(dem_rng_t* r) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  if (r->dummy <= 0 || r->iiy < 0) {
    r->aam = (double)(nextafterf(1.0f, 0.0f) / (float)im);
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

/* PourParticle and InsertParticle (src/MDEM.f90:412-562), as the reference
 * runs them: the largest radius for the height, the spacing test and the
 * wall test; the spacing test on the integer parts of the separations
 * (Rij is declared integer); the inserted particle at rest but for a
 * vertical velocity -1e-4*LengthScale/TimeScale. */
int orc_pour(orc_handle_t h, const dem_pour_t* pr, dem_rng_t* rng, int32_t* inserted) {
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  if (!pr || !rng || !pr->target_number) return fail(h, DEM_ERR_ARG, "null argument");
  if (inserted) *inserted = 0;
  int number[MAXC] = {0}, notreached[MAXC], nnr = 0;
  for (int i = 0; i < h->np; i++) number[h->part[i].index - 1]++;
  for (int c = 0; c < h->nc; c++)
    if (number[c] < pr->target_number[c]) notreached[nnr++] = c + 1;
  if (nnr == 0) return DEM_OK;
  particle_t ip;
  memset(&ip, 0, sizeof ip);
  const int j0 = (int)((double)nnr * orc_ranf(rng)) + 1;
  ip.index = notreached[j0 - 1];
  int ibig = 0;
  for (int c = 1; c < h->nc; c++)
    if (h->radius[c] > h->radius[ibig]) ibig = c;
  const double ri = h->radius[ibig];
  const double di = 2.0 * ri;
  ip.position[2] = h->phalf[2] - ri;
  const double two_pi = 2.0 * 3.14159265358979323846;
  int possible = 0, attempt = 0;
  int* nb = NULL;
  int nbcap = 0;
  while (!possible && attempt < pr->max_attempts) {
    possible = 1;
    attempt++;
    if (pr->shape == DEM_POUR_RECTANGLE) {
      for (int d = 0; d < 2; d++) ip.position[d] = (pr->size[d] - di) * (orc_ranf(rng) - 0.5);
    } else {
      const double rad = 0.5 * orc_ranf(rng) * (pr->size[0] - di);
      const double angle = two_pi * orc_ranf(rng);
      ip.position[0] = rad * sin(angle);
      ip.position[1] = rad * cos(angle);
    }
    const int icell = current_cell(h, ip.position) - 1;
    const int n = mc_neighbour_particles(h, icell, &nb, &nbcap);
    for (int k = 0; possible && k < n; k++) {
      const particle_t* jp = &h->part[nb[k]];
      int32_t s = 0;
      for (int d = 0; d < 3; d++) {
        const int32_t rij = (int32_t)(jp->position[d] - ip.position[d]);
        s += rij * rij;
      }
      possible = (double)s > h->mindijsq[ibig + (jp->index - 1) * h->nc];
    }
    for (int jw = 0; possible && jw < h->nw; jw++) {
      double dij, nij[3], qji[3];
      const int ov = overlap(h, &ip, jw, ri, &dij, nij, qji);
      if (ov < 0) {
        free(nb);
        return fail(h, DEM_ERR_UNSUPPORTED, "pouring position on the axis of cone wall %d", jw + 1);
      }
      possible = !ov;
    }
  }
  free(nb);
  if (!possible) return DEM_OK;
  if (h->np == h->max_np) {
    int m = (int)(1.05 * (double)h->max_np);
    if (m < h->max_np + 1) m = h->max_np + 1;
    resize_variables(h, m);
  }
  ip.velocity[2] = -(1.e-4 * pr->length_scale / pr->time_scale);
  const int k = h->np;
  h->np += 1;
  h->part[k] = ip;
  h->pred[k].index = ip.index;
  memset(&h->prev[(size_t)k * 27], 0, sizeof(double) * 27);
  h->cell_index[k] = current_cell(h, ip.position);
  add_particle(h, k, h->cell_index[k]);
  if (inserted) *inserted = 1;
  return DEM_OK;
}

/* ---- ABI -------------------------------------------------------------------- */

static char g_err[512];

int orc_create(const dem_params_t* p, orc_handle_t* out) {
  if (!p || !out) return DEM_ERR_ARG;
  *out = NULL;
  if (p->nc < 1 || p->nc > MAXC || p->nw < 0 || p->nw > MAXW) {
    snprintf(g_err, sizeof g_err, "nc must be 1..%d and nw 0..%d", MAXC, MAXW);
    return DEM_ERR_ARG;
  }
  if (p->method_order < 1 || p->method_order > 4 || p->n_euler < 1) {
    snprintf(g_err, sizeof g_err, "method_order must be 1..4 and n_euler >= 1");
    return DEM_ERR_ARG;
  }
  orc_handle_t h = (orc_handle_t)calloc(1, sizeof(struct orc_system));
  h->nc = p->nc;
  h->nw = p->nw;
  for (int d = 0; d < 3; d++) {
    h->pbc[d] = p->pbc[d] != 0;
    h->lbox[d] = p->lbox[d];
    h->invl[d] = 1.0 / p->lbox[d]; /* ScaleVariables src/DEM.f90:825 */
    h->phalf[d] = 0.5 * p->lbox[d]; /* InitializeAuxiliaries src/MDEM.f90:118-119 */
    h->mhalf[d] = -h->phalf[d];
    h->g[d] = p->gravity[d];
  }
  h->dt = p->time_step;
  h->order = p->method_order;
  h->nprev = p->method_order - 1;
  h->n_euler = p->n_euler;
  h->next_step = p->method_order;
  const int nc = p->nc, nw = p->nw;
  for (int c = 0; c < nc; c++) {
    h->radius[c] = p->radius[c];
    h->mass[c] = p->mass ? p->mass[c] : 0.0;
    h->inertia[c] = p->inertia ? p->inertia[c] : 0.0;
    h->invm[c] = p->inv_mass[c];
    h->invi[c] = p->inv_inertia[c];
    h->effm[c] = p->eff_mass[c];
  }
  for (int t = 0; t < nc * nc; t++) {
    h->mindij[t] = p->min_dij[t];
    h->mindijsq[t] = p->min_dij_sq[t];
    h->kn[t] = p->kn[t]; h->yn[t] = p->yn[t];
    h->kt[t] = p->kt[t]; h->yt[t] = p->yt[t];
    h->musq[t] = p->musq[t];
  }
  for (int t = 0; t < nc * nw; t++) {
    h->knw[t] = p->knw[t]; h->ynw[t] = p->ynw[t];
    h->ktw[t] = p->ktw[t]; h->ytw[t] = p->ytw[t];
    h->muwsq[t] = p->muwsq[t];
  }
  /* InitializeAuxiliaries (src/MDEM.f90:127-154) */
  double dt = h->dt, aux;
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
  for (int k = 0; k < 3; k++) h->ipc[k] = k + 1; /* InitializeIntegrationVariables */
  int rc = init_cells(h, p->max_ncell > 0 ? p->max_ncell : 30000, p->ref_grid);
  if (rc) {
    snprintf(g_err, sizeof g_err, "%s", h->err);
    orc_destroy(h);
    return rc;
  }
  h->nbr_cap = 256;
  h->nbr_list = (int*)malloc(sizeof(int) * h->nbr_cap);
  *out = h;
  return DEM_OK;
}

static void free_particles(orc_handle_t h) {
  if (h->pairs)
    for (int i = 0; i < h->cap; i++) plist_free(&h->pairs[i]);
  if (h->walls)
    for (int i = 0; i < h->cap; i++) plist_free(&h->walls[i]);
  free(h->pairs); free(h->walls); free(h->part); free(h->pred); free(h->prev);
  free(h->cell_next); free(h->cell_index); free(h->force); free(h->torque);
  free(h->fscale);
  h->pairs = h->walls = NULL;
  h->part = h->pred = NULL;
  h->prev = h->force = h->torque = h->fscale = NULL;
  h->cell_next = h->cell_index = NULL;
  h->np = 0;
  h->cap = 0;
  h->max_np = 0;
  h->has_state = 0;
}

void orc_destroy(orc_handle_t h) {
  if (!h) return;
  free_particles(h);
  free(h->cell_head); free(h->nbr_cells); free(h->n_nbr_cells); free(h->nbr_list);
  free(h->vf); free(h->ef);
  free(h);
}

const char* orc_last_error(orc_handle_t h) { return h ? h->err : g_err; }

int orc_set_walls(orc_handle_t h, int32_t nw, const dem_wall_t* walls) {
  if (nw != h->nw) return fail(h, DEM_ERR_ARG, "nw=%d but the system was created with %d", nw, h->nw);
  for (int j = 0; j < nw; j++) {
    if (walls[j].shape != DEM_SHAPE_RECTANGLE && walls[j].shape != DEM_SHAPE_CONE)
      return fail(h, DEM_ERR_ARG, "wall %d has an invalid shape", j + 1);
    h->wall[j] = walls[j];
  }
  configure_walls(h);
  return DEM_OK;
}

/* WallsInMovement and WM(1:NW) (src/DEM.f90:716-732) */
int orc_set_wall_motion(orc_handle_t h, int32_t nw, const dem_wall_motion_t* m) {
  if (nw != 0 && nw != h->nw)
    return fail(h, DEM_ERR_ARG, "nw=%d but the system was created with %d", nw, h->nw);
  if (nw == 0) {
    h->walls_in_movement = 0;
    return DEM_OK;
  }
  for (int j = 0; j < nw; j++) {
    for (int d = 0; d < 3; d++)
      if (!(m[j].two_pi_frequency[d] >= 0.0))
        return fail(h, DEM_ERR_ARG, "wall %d: two_pi_frequency must be >= 0", j + 1);
    h->wm[j] = m[j];
  }
  h->walls_in_movement = 1;
  return DEM_OK;
}

int orc_download_walls(orc_handle_t h, int32_t nw, dem_wall_t* w) {
  if (nw != h->nw) return fail(h, DEM_ERR_ARG, "nw mismatch");
  for (int j = 0; j < nw; j++) w[j] = h->wall[j];
  return DEM_OK;
}

/* FluidFlow, the specification of ReadFluidSpecifications (scaled) and
 * InitializeFluid (src/mFluid.f90:29-92). */
int orc_set_fluid(orc_handle_t h, const dem_fluid_t* f) {
  free(h->vf);
  free(h->ef);
  h->vf = h->ef = NULL;
  h->fluid = 0;
  if (!f) return DEM_OK;
  if (f->nl < 1 || !f->density) return fail(h, DEM_ERR_ARG, "bad fluid specification");
  h->fluid = 1;
  h->nl = f->nl;
  h->fluid_density = f->fluid_density;
  h->viscosity = f->viscosity;
  h->flow_rate = f->flow_rate;
  for (int c = 0; c < h->nc; c++) h->density[c] = f->density[c];
  h->area = h->lbox[0] * h->lbox[1];
  h->delta_z = h->lbox[2] / (double)h->nl;
  h->vf0 = h->flow_rate / h->area;
  h->h_2 = h->lbox[2] / 2.0;
  h->mh_2 = -h->h_2;
  h->vf = (double*)calloc((size_t)h->nl + 1, sizeof(double));
  h->ef = (double*)calloc((size_t)h->nl + 1, sizeof(double));
  return DEM_OK;
}

int orc_download_fluid_field(orc_handle_t h, int32_t nl, double* vf, double* ef) {
  if (!h->fluid || nl != h->nl) return fail(h, DEM_ERR_ARG, "no fluid with %d layers", nl);
  memcpy(vf, h->vf, sizeof(double) * (size_t)(nl + 1));
  memcpy(ef, h->ef, sizeof(double) * (size_t)(nl + 1));
  return DEM_OK;
}

int orc_set_option(orc_handle_t h, const char* name, double value) {
  if (!strcmp(name, "time")) h->time = value;
  else if (!strcmp(name, "next_step")) h->next_step = (int64_t)value;
  else return fail(h, DEM_ERR_ARG, "unknown option '%s'", name);
  return DEM_OK;
}

typedef struct { int rank, id; } rank_t;
static int cmp_rank(const void* a, const void* b) {
  const rank_t* x = (const rank_t*)a;
  const rank_t* y = (const rank_t*)b;
  if (x->rank != y->rank) return x->rank < y->rank ? -1 : 1;
  return x->id < y->id ? -1 : (x->id > y->id);
}

int orc_upload_state(orc_handle_t h, int32_t np, const int32_t* comp,
                     const double* x, const double* v, const double* w,
                     const double* a, const double* alpha, const double* ring,
                     const int32_t* ipc, const int32_t* cell_rank) {
  if (np < 0 || !comp || !x || !v || !w) return fail(h, DEM_ERR_ARG, "bad state arguments");
  free_particles(h);
  h->np = np;
  h->max_np = np;
  h->cap = np > 0 ? np : 1;
  size_t n = (size_t)h->cap;
  h->part = (particle_t*)calloc(n, sizeof(particle_t));
  h->pred = (particle_t*)calloc(n, sizeof(particle_t));
  h->prev = (double*)calloc(n * 27, sizeof(double));
  h->cell_next = (int*)malloc(sizeof(int) * n);
  h->cell_index = (int*)malloc(sizeof(int) * n);
  h->force = (double*)malloc(sizeof(double) * 3 * n);
  h->torque = (double*)malloc(sizeof(double) * 3 * n);
  h->fscale = (double*)calloc(2 * n, sizeof(double));
  h->pairs = (plist_t*)calloc(n, sizeof(plist_t));
  h->walls = (plist_t*)calloc(n, sizeof(plist_t));
  for (int i = 0; i < np; i++) {
    if (comp[i] < 1 || comp[i] > h->nc) return fail(h, DEM_ERR_ARG, "particle %d has component %d", i + 1, comp[i]);
    particle_t* q = &h->part[i];
    q->index = comp[i];
    h->pred[i].index = comp[i];
    for (int d = 0; d < 3; d++) {
      q->position[d] = x[d * np + i];
      q->velocity[d] = v[d * np + i];
      q->omega[d] = w[d * np + i];
      q->acceleration[d] = a ? a[d * np + i] : 0.0;
      q->alpha[d] = alpha ? alpha[d * np + i] : 0.0;
    }
  }
  if (ring && h->nprev > 0) {
    for (int s = 0; s < h->nprev; s++)
      for (int c = 0; c < 9; c++)
        for (int i = 0; i < np; i++)
          h->prev[((size_t)i * 3 + s) * 9 + c] = ring[((size_t)s * 9 + c) * np + i];
  }
  for (int k = 0; k < 3; k++) h->ipc[k] = k + 1;
  if (ipc)
    for (int k = 0; k < h->nprev; k++) h->ipc[k] = ipc[k];
  /* DistributeParticles3D (src/MCells.f90:542-564), or the given list order. */
  for (int c = 0; c < h->nt; c++) h->cell_head[c] = -1;
  rank_t* order = (rank_t*)malloc(sizeof(rank_t) * n);
  for (int i = 0; i < np; i++) {
    h->cell_index[i] = current_cell(h, h->part[i].position);
    order[i].rank = cell_rank ? cell_rank[i] : 0;
    order[i].id = i;
  }
  if (cell_rank) qsort(order, np, sizeof(rank_t), cmp_rank);
  for (int k = 0; k < np; k++) add_particle(h, order[k].id, h->cell_index[order[k].id]);
  free(order);
  h->has_state = 1;
  return DEM_OK;
}

int orc_download_state(orc_handle_t h, int32_t np, int32_t* comp, double* x,
                       double* v, double* w, double* a, double* alpha,
                       double* ring, int32_t* ipc) {
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np=%d but the state holds %d", np, h->np);
  for (int i = 0; i < np; i++) {
    const particle_t* q = &h->part[i];
    if (comp) comp[i] = q->index;
    for (int d = 0; d < 3; d++) {
      if (x) x[d * np + i] = q->position[d];
      if (v) v[d * np + i] = q->velocity[d];
      if (w) w[d * np + i] = q->omega[d];
      if (a) a[d * np + i] = q->acceleration[d];
      if (alpha) alpha[d * np + i] = q->alpha[d];
    }
  }
  if (ring)
    for (int s = 0; s < h->nprev; s++)
      for (int c = 0; c < 9; c++)
        for (int i = 0; i < np; i++)
          ring[((size_t)s * 9 + c) * np + i] = h->prev[((size_t)i * 3 + s) * 9 + c];
  if (ipc)
    for (int k = 0; k < h->nprev; k++) ipc[k] = h->ipc[k];
  return DEM_OK;
}

int orc_download_predicted(orc_handle_t h, int32_t np, double* x, double* v,
                           double* w, double* a, double* alpha) {
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np mismatch");
  for (int i = 0; i < np; i++) {
    const particle_t* q = &h->pred[i];
    for (int d = 0; d < 3; d++) {
      if (x) x[d * np + i] = q->position[d];
      if (v) v[d * np + i] = q->velocity[d];
      if (w) w[d * np + i] = q->omega[d];
      if (a) a[d * np + i] = q->acceleration[d];
      if (alpha) alpha[d * np + i] = q->alpha[d];
    }
  }
  return DEM_OK;
}

int orc_upload_contacts(orc_handle_t h, int64_t npairs, const uint64_t* keys,
                        const double* ft, const double* vt, int64_t nwall,
                        const uint64_t* wkeys, const double* wft,
                        const double* wvt) {
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  for (int i = 0; i < h->np; i++) {
    h->pairs[i].n = 0;
    h->walls[i].n = 0;
  }
  for (int64_t k = 0; k < npairs; k++) {
    int i = (int)(keys[k] >> 32) - 1, j = (int)(keys[k] & 0xffffffffu) - 1;
    if (i < 0 || j <= i || j >= h->np) return fail(h, DEM_ERR_ARG, "bad pair key %lld", (long long)k);
    plist_t* l = &h->pairs[i];
    int pos = plist_find(l, j);
    if (pos < l->n && l->idx[pos] == j) return fail(h, DEM_ERR_ARG, "duplicate pair key");
    plist_insert(l, pos, j, ft + 3 * k, vt + 3 * k);
  }
  for (int64_t k = 0; k < nwall; k++) {
    int i = (int)(wkeys[k] >> 32) - 1, j = (int)(wkeys[k] & 0xffffffffu) - 1;
    if (i < 0 || i >= h->np || j < 0 || j >= h->nw) return fail(h, DEM_ERR_ARG, "bad wall key");
    plist_t* l = &h->walls[i];
    int pos = plist_find(l, j);
    plist_insert(l, pos, j, wft + 3 * k, wvt + 3 * k);
  }
  return DEM_OK;
}

int orc_download_contacts(orc_handle_t h, int64_t cap, int64_t* npairs,
                          uint64_t* keys, double* ft, double* vt, int64_t wcap,
                          int64_t* nwall, uint64_t* wkeys, double* wft,
                          double* wvt) {
  int64_t n = 0, m = 0;
  /* all MaxNP lists, as SaveContactList writes size(MainParticle) of them
   * (src/MContacts.f90:82-84): after removals some lists past NP remain */
  const int nl = h->max_np > h->np ? h->max_np : h->np;
  for (int i = 0; i < nl; i++) {
    const plist_t* l = &h->pairs[i];
    for (int k = 0; k < l->n; k++, n++) {
      if (n >= cap) continue;
      if (keys) keys[n] = ((uint64_t)(i + 1) << 32) | (uint64_t)(l->idx[k] + 1);
      if (ft) memcpy(ft + 3 * n, l->v1 + 3 * k, sizeof(double) * 3);
      if (vt) memcpy(vt + 3 * n, l->v2 + 3 * k, sizeof(double) * 3);
    }
    const plist_t* lw = &h->walls[i];
    for (int k = 0; k < lw->n; k++, m++) {
      if (m >= wcap) continue;
      if (wkeys) wkeys[m] = ((uint64_t)(i + 1) << 32) | (uint64_t)(lw->idx[k] + 1);
      if (wft) memcpy(wft + 3 * m, lw->v1 + 3 * k, sizeof(double) * 3);
      if (wvt) memcpy(wvt + 3 * m, lw->v2 + 3 * k, sizeof(double) * 3);
    }
  }
  if (npairs) *npairs = n;
  if (nwall) *nwall = m;
  if ((keys && n > cap) || (wkeys && m > wcap))
    return fail(h, DEM_ERR_CAPACITY, "contact arrays too small (%lld pairs, %lld wall)",
                (long long)n, (long long)m);
  return DEM_OK;
}

int orc_download_cells(orc_handle_t h, int32_t np, int32_t* cell_index) {
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np mismatch");
  memcpy(cell_index, h->cell_index, sizeof(int) * np);
  return DEM_OK;
}

int orc_download_cell_rank(orc_handle_t h, int32_t np, int32_t* rank) {
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np mismatch");
  for (int c = 0; c < h->nt; c++) {
    int len = 0;
    for (int q = h->cell_head[c]; q >= 0; q = h->cell_next[q]) len++;
    int r = len;
    for (int q = h->cell_head[c]; q >= 0; q = h->cell_next[q]) rank[q] = --r;
  }
  return DEM_OK;
}

int orc_compute_accelerations_aos(orc_handle_t h, int32_t np, dem_particle_t* p,
                                  double time_step, int32_t upd) {
  if (!h->has_state || np != h->np) return fail(h, DEM_ERR_STATE, "state holds %d particles, call passes %d", h->np, np);
  return compute_accelerations(h, p, time_step, upd != 0);
}

int orc_compute_accelerations(orc_handle_t h, int32_t which, int32_t upd) {
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  return compute_accelerations(h, which ? h->pred : h->part, h->dt, upd != 0);
}

int orc_euler_startup(orc_handle_t h) {
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  return move_particles_euler(h);
}

int orc_step(orc_handle_t h, int64_t nsteps, int64_t first_step,
             int64_t* steps_done, int32_t* escaped) {
  if (!h->has_state) return fail(h, DEM_ERR_STATE, "upload the state first");
  int64_t done = 0;
  int esc = 0;
  for (int64_t s = 0; s < nsteps; s++) {
    int rc = move_particles(h, first_step + s);
    if (rc) return rc;
    done++;
    esc = count_escaped(h);
    if (esc) break;
  }
  h->next_step = first_step + done;
  if (steps_done) *steps_done = done;
  if (escaped) *escaped = esc;
  return esc ? fail(h, DEM_ERR_ESCAPE, "%d particles left the box", esc) : DEM_OK;
}

int orc_stats(orc_handle_t h, dem_stats_t* s) {
  memset(s, 0, sizeof *s);
  s->np = h->np;
  s->pair_contacts = h->pair_contacts;
  s->wall_contacts = h->wall_contacts;
  int64_t stored = 0;
  for (int i = 0; i < h->np; i++) stored += h->pairs[i].n;
  s->stored_pairs = stored;
  s->ncol = h->ncol;
  s->evaluations = h->evaluations;
  s->steps = h->steps;
  s->ref_ncell[0] = h->nx; s->ref_ncell[1] = h->ny; s->ref_ncell[2] = h->nz;
  double ek = 0.0;
  for (int i = 0; i < h->np; i++) {
    const particle_t* q = &h->part[i];
    int c = q->index - 1;
    ek = ek + h->mass[c] * sumsq3(q->velocity) + h->inertia[c] * sumsq3(q->omega);
  }
  s->kinetic_energy = 0.5 * ek;
  s->max_overlap = h->max_overlap;
  s->time = h->time;
  s->next_step = h->next_step;
  return DEM_OK;
}

/* PotentialEnergy (src/MDEM.f90:183-195): Ep = Ep - Mass*(Position .dot.
 * Gravity) over i = 1..NP, the dot product accumulated from zero in index
 * order (dot_product, src/MDimension3D.f90:75-80). */
double orc_potential_energy(orc_handle_t h) {
  double ep = 0.0;
  for (int i = 0; i < h->np; i++) {
    const particle_t* q = &h->part[i];
    double d = 0.0;
    for (int k = 0; k < 3; k++) d = d + q->position[k] * h->g[k];
    ep = ep - h->mass[q->index - 1] * d;
  }
  return ep;
}

int orc_download_force_scale(orc_handle_t h, int32_t np, double* scale) {
  if (np != h->np) return fail(h, DEM_ERR_ARG, "np mismatch");
  for (int i = 0; i < np; i++) {
    scale[i] = h->fscale[2 * i];
    scale[np + i] = h->fscale[2 * i + 1];
  }
  return DEM_OK;
}
