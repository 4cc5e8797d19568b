/*
 * ref_thermal.c - TEST INFRASTRUCTURE ONLY: a call into the reference's own
 * thermal model (src/thermal/thermal_model.c and src/common/grid.c, compiled
 * unmodified by `make -C oracle ref` into _ref/libref_thermal.so) so the tests
 * can pin the oracle's restatement against it.  Nothing ships or measures it.
 */
#include <string.h>

#include "grid.h"
#include "thermal.h"

/* T_out[td*ffd+fd] = the reference's surface temperature for parameters tp
 * (NULL: thermal_params_default) and flux (NULL: its built-in day side). */
int ref_thermal(int ftd, int ffd, const ThermalParams *tp, const double *flux, double *T_out, int nthreads)
{
    SphericalGrid *g = grid_create(ftd, ffd, 1);
    if (!g) return -1;
    ThermalParams p;
    if (tp) p = *tp;
    else thermal_params_default(&p);
    thermal_solve(g, &p, flux, nthreads);
    memcpy(T_out, g->temperature, sizeof(double) * (size_t)ftd * ffd);
    grid_free(g);
    return 0;
}
