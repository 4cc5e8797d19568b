/*
 * common_api.cpp - the `dsmc_common` library API (include/dsmc_common.h) on the
 * host: the reference's integrator, velocity table, grid, collision and
 * polyfit functions (src/common/{integrator,velocity,grid,collision,polyfit}.c)
 * with their names, signatures and struct layouts, so that callers compiled
 * against the reference's headers - its own tests/test_verlet.c among them -
 * link against libdsmc_hip.so alone.
 *
 * Each function is the lane arithmetic the gfx950 kernels run (dsmc_lane.h,
 * dsmc_polyfit.h, dsmc_tables.h) compiled for the host without contraction;
 * where the reference calls libm (exp in collision_check, pow in the grid
 * volumes and heights) or glibc rand_r (velocity_sample), so does this file.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/dsmc_common.h"
#include "dsmc_lane.h"
#include "dsmc_polyfit.h"
#include "dsmc_tables.h"

static_assert(sizeof(dsmc_particle) == 88, "Particle layout (types.h:46-53)");
static_assert(sizeof(dsmc_physics_params) == 136, "PhysicsParams layout (types.h:67-77)");
static_assert(sizeof(dsmc_velocity_table) == 23 * 8 + 23 * 1000 * 8 + 23 * 4 + 4, "VelocityTable layout");
static_assert(sizeof(dsmc_poly_fit) == 56, "PolyFit layout (types.h:89-94, POLY_DEGREE 3)");

extern "C" int verlet_step(dsmc_particle *p, const dsmc_physics_params *params, double h, dsmc_vec3 *oa,
                           double *ds_out)
{
    const double Ri = sqrt(p->pos.x * p->pos.x + p->pos.y * p->pos.y + p->pos.z * p->pos.z);
    const double h2 = h * h;
    double ds = 0.0;
    const double Rn = verlet_push<true>(p->pos.x, p->pos.y, p->pos.z, p->vel.x, p->vel.y, p->vel.z, oa->x, oa->y,
                                        oa->z, Ri, params->G * params->M, h, 1.5 * h2, h2 / 6.0, &ds);
    if (ds_out) *ds_out = ds;
    if (Rn <= params->Radii) return DSMC_VERLET_LANDED;
    if (Rn >= params->max_height) return DSMC_VERLET_ESCAPED;
    return DSMC_VERLET_OK;
}

extern "C" void velocity_table_init(dsmc_velocity_table *vt, double mmass, double mole, double k_B)
{
    const double dT = (240.0 - 20.0) / (DSMC_VT_NT - 1);
    for (int i = 0; i < DSMC_VT_NT; i++) vt->temps[i] = 20.0 + i * dT;
    dsmc::velocity_table(mmass, mole, k_B, &vt->speeds[0][0], vt->count);
}

extern "C" double velocity_sample(const dsmc_velocity_table *vt, double T_local, unsigned int *seed)
{
    const int row = vt_row(T_local);
    return vt->speeds[row][rand_r(seed) % vt->count[row]];
}

extern "C" double velocity_sticking(double vr, double v_old_sq, double alpha)
{
    return sticking_speed(vr, v_old_sq, alpha);
}

extern "C" void grid_spherical_to_index(double thi, double fi, int ftd, int ffd, int *td, int *fd)
{
    sph_index(thi, fi, ftd, ffd, td, fd);
}

/* Volume of every cell of shell [r_lo, r_hi] between latitudes [a, b]: the
 * reference's |r_lo^3 + r_hi^3| / 6 * dphi * |sin a - sin b| (grid.c:51-52,
 * :75-76), the same for every longitude. */
static void shell_volumes(dsmc_sph_grid *g, int ht, double r_lo, double r_hi, double a, double b, int td)
{
    const double dphi = 2.0 * M_PI / g->ffd;
    const double v = fabs(r_lo * r_lo * r_lo + r_hi * r_hi * r_hi) / 6.0 * dphi * fabs(sin(a) - sin(b));
    double *row = g->volume + ((size_t)ht * g->ftd + td) * g->ffd;
    for (int fd = 0; fd < g->ffd; fd++) row[fd] = v;
}

extern "C" void grid_compute_volume_uniform(dsmc_sph_grid *g, double Radii, double unih)
{
    double r_lo = Radii;
    for (int ht = 0; ht < g->hd; ht++) {
        const double r_hi = r_lo + unih;
        double lat_lo = -M_PI / 2.0;
        for (int td = 0; td < g->ftd; td++) {
            const double lat_hi = (td + 1.0) / g->ftd * M_PI - M_PI / 2.0;
            shell_volumes(g, ht, r_lo, r_hi, lat_lo, lat_hi, td);
            lat_lo = lat_hi;
        }
        r_lo = r_hi;
    }
}

extern "C" void grid_compute_volume_exp(dsmc_sph_grid *g, double Radii, double unih, double alpha2)
{
    double r_lo = Radii;
    for (int ht = 0; ht < g->hd; ht++) {
        const double r_hi = r_lo + unih * pow(1.0 + alpha2, ht);
        for (int td = 0; td < g->ftd; td++) {
            const double lat_lo = td > 0 ? (double)td / g->ftd * M_PI - M_PI / 2.0 : -M_PI / 2.0;
            const double lat_hi = (td + 1.0) / g->ftd * M_PI - M_PI / 2.0;
            shell_volumes(g, ht, r_lo, r_hi, lat_lo, lat_hi, td);
        }
        r_lo = r_hi;
    }
}

extern "C" double grid_exp_height(int ht, double alpha2, double unih, double Radii)
{
    return (pow(1.0 + alpha2, ht + 1) - 1.0) / alpha2 * unih + Radii;
}

extern "C" int grid_exp_ht_index(double Ri, double Radii, double unih, double alpha2)
{
    return dsmc::ht_of_r(Ri, Radii, unih, alpha2, 1);
}

extern "C" int collision_check(double *real_col_p, double col_threshold, double ds, double local_density,
                               double cross_section)
{
    *real_col_p += ds * local_density * cross_section;
    return col_threshold < 1.0 - exp(-*real_col_p);
}

extern "C" void collision_execute(dsmc_vec3 *v, dsmc_vec3 v2)
{
    collide_compat(v->x, v->y, v->z, v2.x, v2.y, v2.z);
}

extern "C" double collision_sample_speed(const dsmc_poly_fit *pf, double rand_val, int nn)
{
    if (!pf->valid) return 0.0;
    return sample_speed(pf->coeffs, 3, pf->min_v, pf->max_v, rand_val, nn);
}

extern "C" int polyfit(const double *x, const double *y, int n, int degree, double *coeffs)
{
    return lsq_poly(x, y, n, degree, coeffs);
}

extern "C" double polyval(const double *coeffs, int degree, double x)
{
    return horner(coeffs, degree, x);
}
