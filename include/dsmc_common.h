/*
 * dsmc_common.h - the host-side library API of dsmc_suite's `dsmc_common`
 * (CMakeLists.txt:69-80: src/common/{grid,velocity,integrator,collision,
 * polyfit}.c), exported by libdsmc_hip.so with the reference's symbol names,
 * signatures and struct layouts, so code compiled against the reference's
 * headers (include/{types,integrator,velocity,grid,collision,polyfit}.h) links
 * against this library unchanged.  The file functions of fileio.h and
 * grid_create / grid_free are declared in dsmc_hip.h.
 *
 * These are plain host functions: the same lane arithmetic the gfx950 kernels
 * run (dsmc_lane.h, compiled for the host without contraction), libm where the
 * reference calls libm, glibc rand_r where it draws.  They touch no device.
 */
#ifndef DSMC_COMMON_H
#define DSMC_COMMON_H

#include "dsmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/types.h:7-9 */
typedef struct {
    double x, y, z;
} dsmc_vec3;

/* include/types.h:46-53 (`Particle`); verlet_step reads and writes pos, vel */
typedef struct {
    dsmc_vec3 pos, vel;
    double weight;
    int td, fd, ht;
    double flight_time;
    int collision_count;
} dsmc_particle;

/* include/types.h:67-77 (`PhysicsParams`); verlet_step reads G, M, Radii,
 * max_height */
typedef struct {
    double G, M, k_B, mmass, mole;
    double Radii, ve, lifetime, P;
    double cross_section, alpha;
    double dt_step, dt_rot, max_height, unih, alpha2;
    int use_exp_grid;
} dsmc_physics_params;

/* include/types.h:80-86 (`VelocityTable`): 23 temperatures 20..240 K, up to
 * 1000 tabulated speeds each */
typedef struct {
    double temps[23];
    double speeds[23][1000];
    int count[23];
} dsmc_velocity_table;

/* include/types.h:89-94 (`PolyFit` with POLY_DEGREE 3) */
typedef struct {
    double coeffs[4];
    double min_v, max_v;
    int valid;
} dsmc_poly_fit;

/* include/integrator.h:7-11 (`VerletStatus`) */
enum dsmc_verlet_status {
    DSMC_VERLET_OK = 0,
    DSMC_VERLET_LANDED = 1,   /* |pos| <= Radii after the step */
    DSMC_VERLET_ESCAPED = 2   /* |pos| >= max_height after the step */
};

/* integrator.c:4-55 (integrator.h:19): one Stormer-Verlet step with step h;
 * oa holds the previous acceleration and receives this step's; ds_out (may be
 * NULL) the displacement.  Returns an enum dsmc_verlet_status. */
int verlet_step(dsmc_particle *p, const dsmc_physics_params *params, double h, dsmc_vec3 *oa, double *ds_out);

/* velocity.c:6-47 (velocity.h:7-8): the reference's speed table, with its
 * exp(-3 m v^2 / 2 k T) and 1000-entry truncation per row. */
void velocity_table_init(dsmc_velocity_table *vt, double mmass, double mole, double k_B);
/* velocity.c:49-68 (velocity.h:11-12): a tabulated speed of the row for T,
 * drawn with glibc rand_r(seed). */
double velocity_sample(const dsmc_velocity_table *vt, double T_local, unsigned int *seed);
/* velocity.c:70-73 (velocity.h:15) */
double velocity_sticking(double vr, double v_old_sq, double alpha);

/* grid.c:28-39 (grid.h:16-17): 0-based, clamped */
void grid_spherical_to_index(double thi, double fi, int ftd, int ffd, int *td, int *fd);
/* grid.c:41-60 (grid.h:20): cell volumes of uniform shells, the reference's
 * expression |r_lo^3 + r_hi^3| / 6 dphi |sin th_lo - sin th_hi| kept as it is */
void grid_compute_volume_uniform(dsmc_sph_grid *g, double Radii, double unih);
/* grid.c:62-84 (grid.h:23-24): the same for exponential shells */
void grid_compute_volume_exp(dsmc_sph_grid *g, double Radii, double unih, double alpha2);
/* grid.c:86-89 (grid.h:27): radius at the top of exponential shell ht */
double grid_exp_height(int ht, double alpha2, double unih, double Radii);
/* grid.c:91-99 (grid.h:30) */
int grid_exp_ht_index(double Ri, double Radii, double unih, double alpha2);

/* collision.c:4-13 (collision.h:7-8): accumulates the collision probability;
 * 1 when col_threshold < 1 - exp(-real_col_p) */
int collision_check(double *real_col_p, double col_threshold, double ds, double local_density,
                    double cross_section);
/* collision.c:15-41 (collision.h:13): v <- 1.5 v + 0.5 v2, as the reference
 * evaluates it ((v + v2) + 0.5 (v - v2)) */
void collision_execute(dsmc_vec3 *v, dsmc_vec3 v2);
/* collision.c:43-101 (collision.h:16): bisection of the degree-3 fit for the
 * speed whose value is rand_val (nn - P(max_v)) + P(max_v); 0 for an invalid fit */
double collision_sample_speed(const dsmc_poly_fit *pf, double rand_val, int nn);

/* polyfit.c:10-101 (polyfit.h:10-11): least squares by the normal equations,
 * coeffs[0] the highest power; 0, or -1 (too few points, singular system, or a
 * degree outside 0..6, which the reference does not check) */
int polyfit(const double *x, const double *y, int n, int degree, double *coeffs);
/* polyfit.c:103-110 (polyfit.h:15) */
double polyval(const double *coeffs, int degree, double x);

#ifdef __cplusplus
}
#endif
#endif /* DSMC_COMMON_H */
