/*
 * dem3d_oracle.h -- CPU restatement of the DEM3D per-timestep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the GPU path is
 * compared with; only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  The product (libdem3d_hip.so) never links or calls
 * it.
 *
 * It follows the reference Fortran routine by routine, with the reference's
 * data-structure orders (head-pushed cell lists, reverse-inserted neighbour
 * cell lists, ascending partner lists) and the reference's floating-point
 * operation order, compiled with -ffp-contract=off.  Each function cites the
 * reference lines it restates.  The functions take the same argument structs
 * as include/dem3d.h, prefixed orc_ instead of dem_.
 */
#ifndef DEM3D_ORACLE_H
#define DEM3D_ORACLE_H

#include "../include/dem3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system* orc_handle_t;

int orc_create(const dem_params_t* params, orc_handle_t* out);
void orc_destroy(orc_handle_t h);
const char* orc_last_error(orc_handle_t h);
int orc_set_walls(orc_handle_t h, int32_t nw, const dem_wall_t* walls);
int orc_set_wall_motion(orc_handle_t h, int32_t nw, const dem_wall_motion_t* motion);
int orc_download_walls(orc_handle_t h, int32_t nw, dem_wall_t* walls);
int orc_set_option(orc_handle_t h, const char* name, double value);
double orc_ranf(dem_rng_t* rng);
int orc_set_fluid(orc_handle_t h, const dem_fluid_t* f);
int orc_download_fluid_field(orc_handle_t h, int32_t nl, double* vf, double* ef);
int orc_pour(orc_handle_t h, const dem_pour_t* pour, dem_rng_t* rng, int32_t* inserted);
int orc_remove_escaped(orc_handle_t h, int32_t cap, int32_t* removed, int32_t* nremoved);
int orc_upload_state(orc_handle_t h, int32_t np, const int32_t* comp,
                     const double* x, const double* v, const double* w,
                     const double* a, const double* alpha, const double* ring,
                     const int32_t* ipc, const int32_t* cell_rank);
int orc_download_state(orc_handle_t h, int32_t np, int32_t* comp, double* x,
                       double* v, double* w, double* a, double* alpha,
                       double* ring, int32_t* ipc);
int orc_download_predicted(orc_handle_t h, int32_t np, double* x, double* v,
                           double* w, double* a, double* alpha);
int orc_upload_contacts(orc_handle_t h, int64_t npairs, const uint64_t* keys,
                        const double* ft, const double* vt, int64_t nwall,
                        const uint64_t* wkeys, const double* wft,
                        const double* wvt);
int orc_download_contacts(orc_handle_t h, int64_t cap, int64_t* npairs,
                          uint64_t* keys, double* ft, double* vt, int64_t wcap,
                          int64_t* nwall, uint64_t* wkeys, double* wft,
                          double* wvt);
int orc_download_cells(orc_handle_t h, int32_t np, int32_t* cell_index);
/* Cell-list order of the reference grid: rank[i] such that a larger rank
 * precedes a smaller one inside a cell (the head of each list has the largest
 * rank).  Feeding it back through dem_upload_state(cell_rank=...) reproduces
 * the lists. */
int orc_download_cell_rank(orc_handle_t h, int32_t np, int32_t* rank);
int orc_compute_accelerations_aos(orc_handle_t h, int32_t np,
                                  dem_particle_t* p, double time_step,
                                  int32_t update);
int orc_compute_accelerations(orc_handle_t h, int32_t which, int32_t update);
int orc_euler_startup(orc_handle_t h);
int orc_step(orc_handle_t h, int64_t nsteps, int64_t first_step,
             int64_t* steps_done, int32_t* escaped);
int orc_stats(orc_handle_t h, dem_stats_t* out);
/* PotentialEnergy() (src/MDEM.f90:183-195) of the Particle state. */
double orc_potential_energy(orc_handle_t h);

/* Per-particle |contributions| accumulator used by the force-parity tests
 * (scale_i = |m_eff g| + sum_k (|F_N,ik| + |F_T,ik|), same for torques) of the
 * last evaluation, reference ids, [d*np + i] with d = 0 force, 1 torque. */
int orc_download_force_scale(orc_handle_t h, int32_t np, double* scale);

#ifdef __cplusplus
}
#endif
#endif
