// common.hpp -- structures shared by the host runtime and the gfx950 kernels
// of the DEM per-timestep path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEM_BLOCK 256
#define DEM_MAXW 8

// Per component-pair constants, one 64-byte record per (ci, cj), column-major
// like the reference's NC x NC tables (src/MForces.f90:13-24).
struct PairConst {
  double kn, yn, kt, yt, musq, mindij, mindijsq, rc2;  // rc2 = (mindij+skin)^2
};

// Per component constants (src/MForces.f90:21-24, src/MConfig.f90:32-40).
struct CompConst {
  double radius, inv_mass, inv_inertia, eff_mass, mass, inertia, pad0, pad1;
};

// Per component-wall constants (src/MForces.f90:13-24, Knw..Muw) and the
// per-component cone auxiliaries of ConfigureWalls (src/MConfig.f90:674-682).
struct WallConst {
  double knw, ynw, ktw, ytw, muwsq, max_alpha_u, min_aw_sq, max_aw_sq;
};

// TWall plus the ConfigureWalls auxiliaries of cones (src/MConfig.f90:636-688).
struct WallGeom {
  int shape, pad;
  double l1, l2, l3;
  double u[3], w[3], n[3], pos[3], vel[3], omg[3];
  double rap, sin_t, cos_t;
};

// TWallMovement (src/MDEM.f90:39-45), reduced units.
struct WallMotion {
  double cf[3], amp[3], vamp[3], tpf[3], phase[3];
};

// The fluid of mFluid (src/mFluid.f90:9-23, InitializeFluid :70-92), reduced.
struct FluidConst {
  double area, delta_z, vf0, h_2, mh_2, rho_f, mu;
  int nl, pad;
};

// Device-side control block: flags and counters written by kernels.
struct Ctl {
  int rebuild;          // neighbour list must be rebuilt before the next pass
  int error;            // DEM_KERR_* bits
  int escaped;          // particles outside the box after the last corrector
  int stop;             // set once a step had escapes: later kernels no-op
  int stop_step;        // index (within the dem_step call) of that step
  int max_nnb;          // high-water mark of neighbour-list lengths
  int max_cnt;          // high-water mark of contact-table rows
  int esc_pending;      // escapes found by the running corrector pass
  unsigned long long ncol;          // cumulative (src/MForces.f90:227)
  unsigned long long rebuilds;
  unsigned long long pairs[2];      // pair contacts of the eval, by parity
  unsigned long long wallc[2];      // wall contacts of the eval, by parity
  unsigned long long stored[2];     // stored pair entries after an update eval
  unsigned long long nbpairs;       // directed neighbour entries of last build
  unsigned long long stale_dropped; // stored pairs lost from the list at a rebuild
  unsigned long long pairs_upd_sum; // pair contacts summed over updating passes
  unsigned long long upd_evals;     // updating passes summed
  int max_halo;                     // high-water mark of a tile's halo (staged passes)
  int halo_over;                    // tiles whose halo exceeded DEM_HALO_CAP
};

enum {
  DEM_KERR_NBR_OVERFLOW = 1,
  DEM_KERR_CONTACT_OVERFLOW = 2,
  DEM_KERR_SKIN = 4,
  DEM_KERR_ROLE = 8,
  DEM_KERR_CONE_AXIS = 16,
  DEM_KERR_NAN = 32,
  DEM_KERR_FLUID = 64,  // a particle outside the fluid layers
};

// Constants of one system, passed by value to every kernel.
struct SysConst {
  int np;
  int nc, nw;
  int pbc[3];
  int nprev;
  int hist_mode;
  int max_nb, max_c;
  // reference grid (src/MCells.f90:155-227)
  int rn[3];
  int rc_pack;  // every axis has <= 1024 cells: Kin carries packed coordinates
  double rn_by_l[3], rn_by2[3];
  // neighbour-search grid: gn cells per axis, gb = ceil(gn/8) bricks,
  // gncell = 512 * bricks keys (see grid_key)
  int gn[3];
  int gb[3];
  int gncell;
  int pad1;
  double g_by_l[3], g_by2[3];
  double lbox[3], invl[3], phalf[3];
  double grav[3];
  double skin_half_sq;
  const PairConst* pair;
  const CompConst* comp;
  const WallConst* wallc;
  const WallGeom* walls;
  // neighbour lists with the largest pair cutoff for every pair: set while
  // some Predicted%Index differs from the particle's component (the
  // predictor pass then evaluates it with another radius)
  double rc2_all;
  int cut_all;
};
