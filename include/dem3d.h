/*
 * dem3d.h -- C-ABI of the MI355X (gfx950) DEM per-timestep path.
 *
 * The library replaces, behind plain C types, the module interface that the
 * Fortran driver of DEM3D uses for one time step:
 *
 *   MForces::ComputeAccelerations   src/MForces.f90:145-335
 *   MForces::InitializeForce/CellList/CellDistribution  src/MForces.f90:54-117
 *   MCells  (binning, stencil)      src/MCells.f90:155-227,276-327,498-513,607-643
 *   MContacts (history store)       src/MContacts.f90:329-549
 *   MDEM::MoveParticles             src/MDEM.f90:285-369
 *   MDEM::MoveParticlesEuler        src/MDEM.f90:227-281
 *   MDEM::ApplyPeriodicBoundaryConditions src/MDEM.f90:373-388
 *   MDEM::UpdateParticleList (escape scan only) src/MDEM.f90:612-671
 *   MDEM::KineticEnergy             src/MDEM.f90:199-213
 *
 * Conventions
 *  - All quantities are in the reference's reduced units (src/DEM.f90:807-866).
 *  - Particle ids at the ABI are the reference's 1-based indices; component
 *    indices are 1-based (TParticle%Index).
 *  - NC x NC and NC x NW tables are column-major (Fortran order):
 *    element (i,j), 0-based, lives at i + j*nc.
 *  - SoA particle arrays hold component d of particle i at [d*np + i]
 *    (the layout of a Fortran array X(np,3)).
 *  - The ABM ring is Previous(NPrev,NP) of (Velocity, Acceleration, Alpha):
 *    ring[(s*9 + c)*np + i], s = storage slot 0..nprev-1, c = 0..2 velocity,
 *    3..5 acceleration, 6..8 alpha (src/MDEM.f90:14-24).  ipc[] holds the
 *    reference's 1-based slot order (ipc(1) = most recent).
 *  - Pair keys are (imin << 32) | imax with 1-based ids, ascending; wall keys
 *    are (i << 32) | wall.  ft/vt are 3 doubles per key, exactly the vectors the
 *    reference stores in TPartner%Vector1/Vector2 (src/MContacts.f90:21-26).
 *  - The library owns all device memory; caller arrays are copied and never
 *    retained.  Functions return DEM_OK (0) or a negative DEM_ERR_* code; the
 *    message is available from dem_last_error().
 *  - One handle = one HIP stream; a handle is not thread-safe (like the
 *    reference's module globals).  Distinct handles may run on distinct host
 *    threads.
 */
#ifndef DEM3D_H
#define DEM3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DEM_OK = 0,
  DEM_ERR_ARG = -1,        /* invalid argument or size                       */
  DEM_ERR_DEVICE = -2,     /* HIP runtime failure                             */
  DEM_ERR_CAPACITY = -3,   /* a fixed capacity was exceeded (see message)     */
  DEM_ERR_STATE = -4,      /* call out of order (e.g. step before upload)     */
  DEM_ERR_UNSUPPORTED = -5,/* feature outside the implemented path            */
  DEM_ERR_ESCAPE = -6,     /* particles left the box: host must remove them   */
  DEM_ERR_IO = -7          /* a file could not be read or written, or is not
                              in the expected format                          */
};

enum { DEM_FP64 = 0, DEM_MIXED = 1 };               /* force arithmetic          */
enum { DEM_SHAPE_RECTANGLE = 1, DEM_SHAPE_CONE = 2 }; /* src/MConfig.f90:43-45   */

/* Orientation of stored tangential history.
 * REFERENCE: the stored vectors are the tangential force/velocity of whichever
 *   particle played "i" in the reference's cell loop when they were written,
 *   and are re-read for the current "i" without regard to orientation
 *   (src/MForces.f90:220,262 with src/MContacts.f90:329-423).  The library
 *   reproduces the reference's visitation roles, so forces match it.
 * CANONICAL: the stored vectors always belong to the lower-id particle and are
 *   negated for the other one (physically consistent; differs from the
 *   reference whenever the reference's roles flip). */
enum { DEM_HIST_REFERENCE = 0, DEM_HIST_CANONICAL = 1 };

typedef struct dem_system* dem_handle_t;

/* Constants of one granular system.  Mirrors the module variables of MForces
 * (src/MForces.f90:11-48), MConfig (Lbox, Component) and MDEM (TimeStep,
 * MethodOrder, NumberOfEulerIntervals) after InitializeForce and
 * ScaleVariables have run. */
typedef struct {
  int32_t nc;               /* number of components (1..DEM_MAX_COMPONENTS)   */
  int32_t nw;               /* number of walls (0..DEM_MAX_WALLS)             */
  int32_t pbc[3];           /* PBC(dim) (nonzero = periodic)                  */
  int32_t method_order;     /* MethodOrder, 1..4                              */
  int32_t n_euler;          /* NumberOfEulerIntervals (>= 1)                  */
  int32_t max_ncell;        /* MaxNCell of the reference grid (30000)         */
  int32_t ref_grid[3];      /* reference grid Nx,Ny,Nz; all 0 = derive as
                               InitCellStructure3D does (src/MCells.f90:155)  */
  int32_t precision;        /* DEM_FP64 or DEM_MIXED                          */
  int32_t history_mode;     /* DEM_HIST_REFERENCE or DEM_HIST_CANONICAL       */
  int32_t max_neighbours;   /* neighbour-list capacity per particle (1..64), 0 = 48 */
  int32_t max_contacts;     /* contact-table capacity per particle, 0 = 24    */
  double lbox[3];           /* Lbox                                            */
  double gravity[3];        /* Gravity                                         */
  double time_step;         /* TimeStep                                        */
  double skin;              /* neighbour-list skin; <= 0 selects 0.1*min diam  */
  const double* radius;     /* [nc]  Component%Radius                          */
  const double* mass;       /* [nc]  Component%Mass (energies)                 */
  const double* inertia;    /* [nc]  Component%Inertia (energies)              */
  const double* inv_mass;   /* [nc]  InvMass                                   */
  const double* inv_inertia;/* [nc]  InvInertia                                */
  const double* eff_mass;   /* [nc]  EffectiveMass                             */
  const double* min_dij;    /* [nc*nc] MinDij                                  */
  const double* min_dij_sq; /* [nc*nc] MinDijSq                                */
  const double* kn;         /* [nc*nc] Kn  */
  const double* yn;         /* [nc*nc] Yn  */
  const double* kt;         /* [nc*nc] Kt  */
  const double* yt;         /* [nc*nc] Yt  */
  const double* musq;       /* [nc*nc] MuSq */
  const double* knw;        /* [nc*nw] Knw */
  const double* ynw;        /* [nc*nw] Ynw */
  const double* ktw;        /* [nc*nw] Ktw */
  const double* ytw;        /* [nc*nw] Ytw */
  const double* muwsq;      /* [nc*nw] MuwSq */
} dem_params_t;

#define DEM_MAX_COMPONENTS 64
#define DEM_MAX_WALLS 8

/* TWall (src/MDimension3D.f90:45-59), lengths already halved as
 * ReadSpecifications does (src/DEM.f90:477-479). */
typedef struct {
  int32_t shape;            /* DEM_SHAPE_*                                     */
  int32_t order;            /* ignored: recomputed as ConfigureWalls does      */
  double l1, l2, l3;
  double stiffness, shear_stiffness;   /* informational (tables carry Knw..)  */
  double u[3], w[3], n[3];
  double position[3], velocity[3], omega[3];
} dem_wall_t;

/* TWallMovement (src/MDEM.f90:39-45), one per wall, in reduced units (as
 * ReadSpecifications and ScaleVariables leave it, src/DEM.f90:719-732 and
 * :841-846).  MoveWalls(Time) (src/MDEM.f90:392-408) sets, per axis,
 *   Phase    = two_pi_frequency*Time + phase_angle, reduced by 2 pi
 *   Position = constant_factor + amplitude*cos(Phase)
 *   Velocity = velocity_amplitude*sin(Phase). */
typedef struct {
  double constant_factor[3];
  double amplitude[3];
  double velocity_amplitude[3];
  double two_pi_frequency[3];
  double phase_angle[3];
} dem_wall_motion_t;

/* TParticle (src/MDimension3D.f90:28-35) byte layout: 128 B records. */
typedef struct {
  int32_t index;            /* component, 1-based */
  int32_t pad_;
  double position[3], velocity[3], acceleration[3], omega[3], alpha[3];
} dem_particle_t;

/* Counters and observables of the last evaluation / step. */
typedef struct {
  int64_t np;               /* particles                                       */
  int64_t pair_contacts;    /* P: overlapping particle pairs (i<j)             */
  int64_t wall_contacts;    /* W: overlapping particle-wall pairs              */
  int64_t stored_pairs;     /* entries in the pair history (incl. stale ones)  */
  int64_t ncol;             /* Ncol (src/MForces.f90:47,227)                   */
  int64_t evaluations;      /* force evaluations done                          */
  int64_t steps;            /* MoveParticles steps done                        */
  int64_t rebuilds;         /* neighbour-list rebuilds                         */
  int64_t neighbour_pairs;  /* directed entries in the neighbour list          */
  int64_t stale_dropped;    /* stale pairs lost: stale-pair table (4096) full  */
  int64_t ref_ncell[3];     /* reference grid                                  */
  int64_t nb_ncell[3];      /* neighbour-search grid                           */
  double kinetic_energy;    /* KineticEnergy() (src/MDEM.f90:199-213)          */
  double max_overlap;       /* max (MinDij - Dij)/MinDij over pair contacts    */
  double ms_pass1;          /* mean device time of the last dem_step, per call */
  double ms_pass2;
  double ms_rebuild;        /* mean over the rebuilds that ran                 */
  double ms_step;           /* mean device time per step                       */
  double mean_pair_contacts;/* P averaged over the first passes of the last
                               dem_step call                                   */
  double time;              /* Time (src/MConfig.f90:61) of the last MoveWalls
                               point reached: step*TimeStep after a step,
                               advanced by TimeStep/NumberOfEulerIntervals per
                               Euler sub-step                                  */
  int64_t next_step;        /* step number a batch gives the system's next
                               MoveParticles (first_step + steps of the last
                               dem_step; MethodOrder initially)                */
  int64_t max_halo;         /* largest halo of a 256-particle tile of the
                               neighbour lists built so far (staged passes)    */
  int64_t halo_overflow;    /* tiles whose halo exceeded its 512 staged rows
                               (their outer partners are read from memory)     */
  int64_t device_bytes;     /* device memory the handle holds for its state and
                               constants (dem_footprint of its parameters and
                               particle count); buffers made on demand (the
                               observables' partial sums, the level-1 drop-in's
                               staging, the fluid layers) are not counted      */
} dem_stats_t;

/* ---- lifetime ------------------------------------------------------------ */
int dem_create(const dem_params_t* params, dem_handle_t* out);
/* No counterpart in the reference.  The device bytes dem_create and a state
 * of np particles allocate for params (the figure dem_stats reports as
 * device_bytes once such a state is uploaded), computed on the host without
 * a device: what a request loop admits by. */
int dem_footprint(const dem_params_t* params, int32_t np, int64_t* bytes);
void dem_destroy(dem_handle_t h);
/* Message of the last error on h (h may be NULL for dem_create failures). */
const char* dem_last_error(dem_handle_t h);

/* ---- state in / out ------------------------------------------------------ */
/* Walls; replaces Wall(:) (src/MConfig.f90:70) and runs ConfigureWalls. */
int dem_set_walls(dem_handle_t h, int32_t nw, const dem_wall_t* walls);
/* WallsInMovement = .true. with WM(1:nw) = motion (src/DEM.f90:716-732);
 * nw = 0 (motion ignored) sets WallsInMovement = .false.  The walls then move
 * where the reference calls MoveWalls: before the corrector pass of every
 * step, at Time = step*TimeStep (src/MDEM.f90:333-337), and after every Euler
 * sub-step, at the accumulated Time (src/MDEM.f90:273-277).  Until the first
 * move they keep the positions and velocities of dem_set_walls. */
int dem_set_wall_motion(dem_handle_t h, int32_t nw, const dem_wall_motion_t* motion);
/* Wall(1:nw) as they stand now (positions and velocities moved). */
int dem_download_walls(dem_handle_t h, int32_t nw, dem_wall_t* walls);

/* Particle state.  a/alpha/ring/ipc/cell_rank may be NULL (zeros; fresh ring;
 * fresh cell lists as DistributeParticles3D builds them).  cell_rank gives the
 * order of the reference's cell lists: within one cell a particle with a larger
 * rank precedes one with a smaller rank (ties: larger id first).  Right after
 * dem_euler_startup, until the next evaluation, the lists still hold the
 * particles that the last Euler sub-step moved into another cell in their old
 * cell; a state downloaded then (or written by dem_save_state_file then) does
 * not carry that, and a pair of such a particle in its new cell can resume
 * with the other role.  States taken after a MoveParticles step resume within
 * rounding (the rebuilt lists sum in another order). */
int dem_upload_state(dem_handle_t h, int32_t np, const int32_t* comp,
                     const double* x, const double* v, const double* w,
                     const double* a, const double* alpha,
                     const double* ring, const int32_t* ipc,
                     const int32_t* cell_rank);
int dem_download_state(dem_handle_t h, int32_t np, int32_t* comp,
                       double* x, double* v, double* w,
                       double* a, double* alpha, double* ring, int32_t* ipc);
/* Predicted(:) of the last MoveParticles (src/MDEM.f90:20). */
int dem_download_predicted(dem_handle_t h, int32_t np, double* x, double* v,
                           double* w, double* a, double* alpha);

/* Contact history (MContacts lists).  Keys ascending, 1-based. */
int dem_upload_contacts(dem_handle_t h, int64_t npairs, const uint64_t* keys,
                        const double* ft, const double* vt, int64_t nwall,
                        const uint64_t* wkeys, const double* wft,
                        const double* wvt);
/* Any output pointer may be NULL to query counts only. */
int dem_download_contacts(dem_handle_t h, int64_t cap, int64_t* npairs,
                          uint64_t* keys, double* ft, double* vt, int64_t wcap,
                          int64_t* nwall, uint64_t* wkeys, double* wft,
                          double* wvt);
/* CellIndex(NP) (1-based) of the reference grid after the last evaluation. */
int dem_download_cells(dem_handle_t h, int32_t np, int32_t* cell_index);

/* ---- the hot path -------------------------------------------------------- */
/* Drop-in for ComputeAccelerations(Particle, NP, TimeStep, Update)
 * (src/MForces.f90:145): reads Index/Position/Velocity/Omega of p[0..np),
 * writes Acceleration and Alpha.  Contact history and cell order stay resident. */
int dem_compute_accelerations_aos(dem_handle_t h, int32_t np, dem_particle_t* p,
                                  double time_step, int32_t update);
/* Same on the resident state: which = 0 Particle, 1 Predicted. */
int dem_compute_accelerations(dem_handle_t h, int32_t which, int32_t update);
/* One MoveParticlesEuler (src/MDEM.f90:258-281) on the resident state. */
int dem_euler_startup(dem_handle_t h);
/* nsteps calls of MoveParticles(step) (src/MDEM.f90:285-369) followed each by
 * the escape scan of UpdateParticleList.  Stops after the first step in which
 * a particle left the box (returns DEM_ERR_ESCAPE, *escaped > 0).
 * steps_done / escaped may be NULL. */
int dem_step(dem_handle_t h, int64_t nsteps, int64_t first_step,
             int64_t* steps_done, int32_t* escaped);
/* Waits for the handle's stream. */
int dem_synchronize(dem_handle_t h);
int dem_stats(dem_handle_t h, dem_stats_t* out);

/* Runtime options (no counterpart in the reference):
 *   "timing"            1 = time pass 1, pass 2 and rebuild kernels with HIP
 *                       events inside dem_step (reported by dem_stats)
 *   "resort_interval"   steps between re-sorts of the particle arrays into
 *                       cell order (0 = only at upload); default 500
 *   "keep_predicted_acc" 1 = keep Predicted%Acceleration/Alpha for
 *                       dem_download_predicted (one extra 48 B write per
 *                       particle and step); default 0
 *   "time"              sets Time (src/MConfig.f90:61), e.g. from a saved
 *                       configuration before an Euler start-up
 *   "next_step"         sets the step number of the next batched step
 *   "pass_shape"        launch shape of the force passes (see dem_tune_t)
 *   "block_map"         block order of the force passes (see dem_tune_t)
 *   "halo_rows"         rows of a tile's halo the LDS-staged launch shapes
 *                       stage (1..512, default 512); neighbours beyond them are
 *                       read from memory (the results do not change) */
int dem_set_option(dem_handle_t h, const char* name, double value);

/* ---- observables ---------------------------------------------------------- */
/* Bulk observables of the resident Particle state, computed on the GPU with
 * fixed-order reductions (bitwise reproducible).  The energies are the
 * reference's; the rest is defined here (the reference driver has no
 * coordination number or packing fraction). */
typedef struct {
  double kinetic_energy;    /* KineticEnergy() (src/MDEM.f90:199-213)          */
  double potential_energy;  /* PotentialEnergy() (src/MDEM.f90:183-195):
                               -sum Mass*(Position . Gravity)                   */
  int64_t pair_contacts;    /* overlapping pairs at the current positions
                               (MinDijSq - DijSq > 0, each pair once)           */
  int64_t wall_contacts;    /* overlapping particle-wall pairs (Overlap)       */
  double coordination;      /* 2 * pair_contacts / NP                          */
  double packing_fraction;  /* solid_volume / (Lx * Ly * (region[1] - region[0])) */
  double region[2];         /* [zlo, zhi): the whole box when z is periodic,
                               else from one largest diameter above -Lz/2 to
                               one largest diameter below `top`               */
  double solid_volume;      /* volume of the spheres whose centre lies in the
                               region                                          */
  double top;               /* max (z + r) over the particles                  */
} dem_observables_t;
/* Observables of the Particle state.  With nlayers > 0, profile[4*nlayers]
 * receives, for nlayers equal layers of the box height from -Lz/2, the
 * volume of sphere material inside each layer (exact sphere-slab
 * intersection) over the layer volume, then the volume-weighted mean
 * velocity (x, y, z) of that material (zero in an empty layer).  Device
 * scratch for the partial sums is allocated on first use and kept by the
 * handle; nlayers <= 65536.  Waits for the handle's stream. */
int dem_observables(dem_handle_t h, int32_t nlayers, dem_observables_t* out, double* profile);
/* dem_observables without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, code that must not call hipMalloc).  The partial sums live in one
 * fixed array of the library in each GPU's memory (DEM_OBS_NOALLOC_BYTES,
 * about 2.1 MB, held until the process ends), so nlayers <= DEM_OBS_MAX_LAYERS;
 * larger values return DEM_ERR_CAPACITY.  The array is shared by every handle
 * on that GPU: calls from one process are serialised by the library, calls
 * that overlap on one GPU from other code paths would race.  Results are
 * bitwise those of dem_observables. */
#define DEM_OBS_MAX_LAYERS 256
#define DEM_OBS_NOALLOC_BYTES ((256 * 8 + 256 * DEM_OBS_MAX_LAYERS * 4 + 16 + 4 * DEM_OBS_MAX_LAYERS) * 8)
int dem_observables_noalloc(dem_handle_t h, int32_t nlayers, dem_observables_t* out,
                            double* profile);

/* ---- fluid drag (mFluid) --------------------------------------------------- */
/* An upward fluid flow through the box (src/mFluid.f90), reduced units.  At
 * every force evaluation the porosity Ef and the fluid velocity Vf of NL+1
 * horizontal layers are computed from the evaluated particles
 * (UpdateVelocityField, :96-140), and every particle feels the drag of
 * DragForce (:144-197) at the layer values interpolated to its height.  The
 * buoyancy is in dem_params_t.eff_mass (EffectiveMass = m(rho - rho_f)/rho). */
typedef struct {
  double fluid_density;     /* FluidDensity                                    */
  double viscosity;         /* FluidViscosity                                  */
  double flow_rate;         /* FluidFlowRate (volume per time, upward)         */
  int32_t nl;               /* NL: layers 0..NL over the box height (>= 1)     */
  int32_t pad_;
  const double* density;    /* [nc] Component%Density                          */
} dem_fluid_t;
/* FluidFlow = .true. with f; f = NULL turns it off. */
int dem_set_fluid(dem_handle_t h, const dem_fluid_t* f);
/* Vf(0:NL) and Ef(0:NL) of the last evaluation. */
int dem_download_fluid_field(dem_handle_t h, int32_t nl, double* vf, double* ef);

/* ---- particle number changes --------------------------------------------- */
/* State of the reference's random-number generator (MUtil: aam, Dummy, iix,
 * iiy, iik).  A run starts with dummy = the (negative) seed of the
 * specification and iix = iiy = -1. */
typedef struct {
  double aam;
  int32_t dummy, iix, iiy, iik;
} dem_rng_t;
/* ranf(Dummy) (src/MUtil.f90:16-44): the next deviate in (0, 1). */
double dem_ranf(dem_rng_t* rng);

#define DEM_POUR_RECTANGLE 1
#define DEM_POUR_CIRCLE 2
/* A pouring specification (the .prg file, src/DEM.f90:749-803), reduced. */
typedef struct {
  const int32_t* target_number;  /* [nc] TargetNumber                         */
  int32_t shape;                 /* DEM_POUR_RECTANGLE or DEM_POUR_CIRCLE     */
  int32_t max_attempts;          /* MaxAttempts                               */
  double size[2];                /* PouringSize (1 value for a circle)        */
  double length_scale, time_scale;  /* for the inserted velocity            */
} dem_pour_t;
/* One PourParticle (src/MDEM.f90:412-536): picks a component short of its
 * target and up to max_attempts positions on the pouring surface at the top
 * of the box, and inserts the particle (InsertParticle, :540-562) as number
 * NP+1 at the head of its cell list, when none of the tests fails.  rng
 * advances as the reference's generator does.  *inserted = 1 when a particle
 * was added.  The driver calls it before the steps with mod(Step, IPouring)
 * == 0 (src/DEM.f90:115-123). */
int dem_pour(dem_handle_t h, const dem_pour_t* pour, dem_rng_t* rng, int32_t* inserted);
/* UpdateParticleList (src/MDEM.f90:612-671) on the resident state: every
 * particle outside the box is removed, from the last index down, the last
 * particle taking its place (cell-list position, ABM ring and contact
 * lists renamed as the reference renames them).  removed[0..cap) receives
 * the 1-based index each removal freed, in order; *nremoved the count.
 * Call it after dem_step returned DEM_ERR_ESCAPE, then step on. */
int dem_remove_escaped(dem_handle_t h, int32_t cap, int32_t* removed, int32_t* nremoved);

/* ---- files in the reference's binary formats ----------------------------
 * BinFiles byte streams (src/BinFiles.f90): every value as its raw bytes in
 * the order written, strings without a length, no record markers.  The
 * functions without a handle work on host arrays and need no device.  Their
 * errors are read with dem_last_error(NULL). */
#define DEM_CFG_COMPLETE 72   /* Complete (src/MConfig.f90:20) */
#define DEM_CFG_COMPACT 28    /* Compact */

/* A configuration: what SaveConfiguration writes (src/MConfig.f90:109-236)
 * and ReadConfiguration reads (:266-448).  Arrays belong to the caller.
 * dem_config_read sets form, version, the scalars and nc, nw, np, then fills
 * every array that is not NULL, so a call with NULL arrays queries the sizes.
 * The Compact form stores positions as 16-bit integers and wall directions as
 * two 16-bit angles; it has no mass scale, component properties beyond
 * Number and Radius, wall velocities or particle velocities (those read back
 * as zeros).  Files with orientation vectors (Direction) are read with the
 * vectors skipped and written without them. */
typedef struct {
  int32_t form;             /* DEM_CFG_COMPLETE or DEM_CFG_COMPACT            */
  int32_t version;          /* file version: 2 when written; 1 or 2 read      */
  double length_scale, time_scale, mass_scale;
  double time;              /* Time                                            */
  double lbox[3];
  int32_t nc, nw, np;
  int32_t* number;          /* [nc] Component%Number (counts of each comp.)   */
  double* radius;           /* [nc] Component%Radius ... %Inertia             */
  double* density;
  double* stiffness;
  double* shear_stiffness;
  double* mass;
  double* inertia;
  dem_wall_t* walls;        /* [nw] Wall(:) with lengths halved               */
  int32_t* comp;            /* [np] Particle%Index (1-based)                  */
  double* x;                /* [3][np] Position                               */
  double* v;                /* [3][np] Velocity                               */
  double* w;                /* [3][np] Omega                                  */
} dem_config_t;
int dem_config_write(const char* path, const dem_config_t* cfg);
int dem_config_read(const char* path, dem_config_t* cfg);

/* The driver's own records of a state file (SaveState, src/DEM.f90:947-1014):
 * run bookkeeping before the configuration, random-number state and step
 * counters after the contact list. */
typedef struct {
  int32_t finished;         /* Finished (2-byte logical: 1 = true)            */
  char start_time[28];      /* StartTime                                       */
  int32_t n_interruptions;  /* NumberOfInterruptions                           */
  char* interruption;       /* [28*n] Interruption(:); read: NULL to query    */
  char* retaking;           /* [28*n] Retaking(:)                              */
  char finish_time[28];     /* FinishTime                                      */
  int32_t hours, minutes, seconds, milisecs;  /* Duration (RealTime)        */
  double aam;               /* random-number generator state (MUtil):        */
  int32_t dummy, iix, iiy, iik;   /* aam, Dummy, iix, iiy, iik               */
  int32_t conf;             /* Conf                                            */
  int32_t next_step;        /* Step + 1                                        */
  int32_t ncol;             /* Ncol                                            */
} dem_run_t;

/* The hot path's records of a state file: the cell lists
 * (SaveCellStructure, src/MCells.f90:960-983), the ABM ring
 * (SaveIntegrationVariables, src/MDEM.f90:52-68) and the contact lists
 * (SaveContactList, src/MContacts.f90:75-118).  Reading fills what is not
 * NULL after setting the counts. */
typedef struct {
  int32_t nprev;            /* NPrev (MethodOrder - 1); must be set to read   */
  int32_t ipc[3];
  double* ring;             /* [nprev][9][np], as dem_upload_state            */
  int32_t ncell;            /* Nt                                              */
  int32_t* cell_len;        /* [ncell] particles in each cell list            */
  int32_t* cell_members;    /* [np] the lists one after another, list order   */
  int64_t npairs;           /* pair contacts, keys (imin<<32)|imax ascending  */
  uint64_t* keys;
  double* ft;               /* [npairs][3] Vector1                             */
  double* vt;               /* [npairs][3] Vector2                             */
  int64_t nwall;            /* wall contacts, keys (i<<32)|wall ascending     */
  uint64_t* wkeys;
  double* wft;
  double* wvt;
} dem_state_sections_t;
int dem_state_file_write(const char* path, const dem_run_t* run, const dem_config_t* cfg,
                         const dem_state_sections_t* sec);
/* cfg and sec are filled as by dem_config_read (Complete form); sec->nprev
 * must be set by the caller (the file does not record it). */
int dem_state_file_read(const char* path, dem_run_t* run, dem_config_t* cfg,
                        dem_state_sections_t* sec);

/* The state of a handle as a state file: particles, walls, Time, cell
 * lists, ABM ring and contacts come from the device; meta gives what the
 * library does not hold (scales, Component Density and stiffnesses; its
 * array members other than those, and its particle arrays, are ignored;
 * Component%Number is counted from the particles). */
int dem_save_state_file(dem_handle_t h, const char* path, const dem_run_t* run,
                        const dem_config_t* meta);
/* Resume from a state file (ReadState, src/DEM.f90:1016-1110): uploads the
 * particles (accelerations zero: the next MoveParticles recomputes them),
 * ring, cell-list order and contacts, sets Time and next_step = Step + 1.
 * The file's components and walls must be those of the handle (nc, nw).
 * run and meta may be NULL; meta's arrays are filled where not NULL. */
int dem_load_state_file(dem_handle_t h, const char* path, dem_run_t* run, dem_config_t* meta);
/* Cell-list order of the last evaluation: within each cell, the particle
 * with the largest rank heads the list (ranks 0..n-1 per cell). */
int dem_download_cell_rank(dem_handle_t h, int32_t np, int32_t* rank);

/* ---- batches ------------------------------------------------------------
 * No counterpart in the reference (it runs one system per process).  For
 * callers that run many independent systems: a batch steps them together,
 * every phase of a step (the two force passes, their counter sums, each
 * kernel of the neighbour-list rebuild) being one launch for all of its
 * systems.  Systems are created, uploaded and started (dem_euler_startup)
 * as usual and then added; while in a batch they are stepped only by
 * dem_batch_step (dem_step on them fails), and downloads and dem_stats work
 * between batch steps.  All systems of a batch use one precision mode;
 * batched systems are not re-sorted. */
typedef struct dem_batch* dem_batch_t;
int dem_batch_create(dem_batch_t* out);
void dem_batch_destroy(dem_batch_t b);  /* leaves the systems alive */
const char* dem_batch_last_error(dem_batch_t b);
int dem_batch_add(dem_batch_t b, dem_handle_t h);
int dem_batch_remove(dem_batch_t b, dem_handle_t h);
int dem_batch_size(dem_batch_t b);
/* nsteps MoveParticles steps of every system; steps_done[i] / escaped[i]
 * (may be NULL) per system in the order of addition.  A system in which
 * particles escaped stops after that step (DEM_ERR_ESCAPE is returned).
 * The steps of a system are numbered from its next_step (dem_stats), which
 * sets Time for its moving walls. */
int dem_batch_step(dem_batch_t b, int64_t nsteps, int64_t* steps_done,
                   int32_t* escaped);
/* "timing": 1 = record the device time of every step of a dem_batch_step
 * call, read back with dem_batch_step_times (returns the count).
 * "graph": 1 = replay the step's launches from HIP graphs captured once per
 * batch layout (recaptured when systems join or leave); 0 (default) = launch
 * them one by one (as fast on this hardware, and no recapture when the
 * membership changes).  "graph_steps": steps per replayed graph (1..256,
 * default 8); with "timing" every step is its own replay.
 * "pass_shape", "block_map": launch shape and block order of the force
 * passes (see dem_tune_t).
 * "persistent": 1 = step the batch with persistent cooperative launches:
 * every system is stepped by its own group of resident blocks (sized by its
 * particles), the groups synchronise only their own blocks between the
 * phases of a step, and one launch runs "persist_steps" steps (1..100000,
 * default 250) of every system; results are bitwise those of the
 * launch-per-phase path.  Needs one resident block per system (else
 * DEM_ERR_CAPACITY); "timing" uses the launch-per-phase path.
 * "spin_limit": bound of a group barrier's wait (polls, default 2^24); a
 * wait that runs out ends the launch and the call returns DEM_ERR_DEVICE
 * (the systems' states are then not valid). */
int dem_batch_set_option(dem_batch_t b, const char* name, double value);
int64_t dem_batch_step_times(dem_batch_t b, int64_t cap, double* ms);

/* ---- launch shapes and the autotuner ------------------------------------ */
/* No counterpart in the reference.  The force passes come in DEM_NSHAPES
 * launch shapes: lanes per block, the occupancy (wavefronts per SIMD) their
 * register allocation targets, and a variant: bit 0, whether the partner's
 * 64-byte velocity record is gathered with its position for every list entry
 * or only for the entries the list rebuild put first as contacts (the others
 * fetch it when they overlap); bit 1, whether each block first stages in LDS
 * the records of its tile of 256 particles and the positions of the tile's
 * halo (its neighbours outside the tile, listed by every neighbour-list
 * rebuild), so that most partner records come from LDS.  Every shape gives
 * bitwise the same results, so a shape
 * changes speed only; shape 0 (256 lanes, 3 waves, contacts only) is the
 * default.  The autotuner times every shape on the live system while
 * stepping it, over steps_per_shape steps timed with HIP events: half of
 * them in a sweep over the shapes in order and the rest in a sweep back
 * (the first launch of a shape, which loads its code, is an untimed step
 * before its first timed steps), so that a trend in the system (contacts
 * forming or relaxing) weighs on every shape alike; the fastest becomes the
 * system's (batch's) shape.  Then the best shape is timed in both block
 * orders, current and other, in the order A B B A with the same halves (one
 * untimed step before the other order's first steps), and the faster order
 * is kept: "block_map" 0 gives each XCD consecutive particle ranges, 1
 * keeps the dispatch order (all XCDs in one region of the particles at a
 * time).  For s = steps_per_shape the tuning takes
 * DEM_NSHAPES * (s + 1) + 2 * s + 1 steps.  They count as ordinary
 * MoveParticles steps (numbered from first_step, resp. each system's
 * next_step), and the trajectory is bitwise the one dem_step
 * (dem_batch_step) gives for the same number of steps.  Escapes and errors stop the
 * tuning with dem_step's return code, the shape and order unchanged.  Both
 * can also be set directly with the options "pass_shape" and "block_map". */
#define DEM_NSHAPES 8
typedef struct {
  int32_t nshapes;             /* shapes timed                                 */
  int32_t chosen;              /* the fastest, now in use                      */
  int64_t steps;               /* MoveParticles steps the tuning advanced      */
  int32_t block[DEM_NSHAPES];  /* lanes per block of each shape                */
  int32_t waves[DEM_NSHAPES];  /* wavefronts per SIMD it targets               */
  int32_t front[DEM_NSHAPES];  /* variant: bit 0 set, partner velocities
                                  gathered with the position only for the
                                  contact-first entries of a list (clear: for
                                  every entry); bit 1, tile and halo staged
                                  in LDS                                       */
  double ms[DEM_NSHAPES];      /* mean device ms per step: the two passes
                                  (dem_autotune), the whole step
                                  (dem_batch_autotune)                         */
  int32_t block_map;           /* block order chosen (option "block_map")      */
  int32_t pad_;
  double ms_map[2];            /* the chosen shape in block order 0 and 1      */
} dem_tune_t;
int dem_autotune(dem_handle_t h, int64_t steps_per_shape, int64_t first_step,
                 dem_tune_t* out);
int dem_batch_autotune(dem_batch_t b, int64_t steps_per_shape, dem_tune_t* out);

#ifdef __cplusplus
}
#endif
#endif /* DEM3D_H */
