/*
 * dsmc_tables.h - host-side construction of everything a run's kernels read:
 * derived scalars (computed the way the reference computes them), the velocity
 * table, exact cell-index threshold tables, the icosphere hierarchy with its edge
 * normals and launch CDF.  Host C++ only (libm); shared by the HIP library and the
 * test-only host build of the lane state machine.
 */
#ifndef DSMC_TABLES_H
#define DSMC_TABLES_H

#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>

#include "../../include/dsmc_hip.h"
#include "dsmc_lane.h"

namespace dsmc {

/* Order-preserving map double <-> uint64 (non-NaN), for bisection over doubles. */
inline uint64_t okey(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
inline double ofrom(uint64_t k)
{
    uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double x;
    memcpy(&x, &u, 8);
    return x;
}

/* Smallest double x in [lo, hi] with f(x) >= k, given f(lo) < k <= f(hi) and f
 * non-decreasing.  Returns NAN if the bracket is invalid or f is not monotone in
 * a window of +-window ulps around the threshold. */
template <class F>
double first_ge(F f, int k, double lo, double hi, int window = 256)
{
    uint64_t a = okey(lo), b = okey(hi);
    if (!(f(lo) < k) || !(f(hi) >= k)) return NAN;
    while (b - a > 1) {
        uint64_t m = a + (b - a) / 2;
        if (f(ofrom(m)) >= k) b = m; else a = m;
    }
    for (int i = 1; i <= window; i++) {
        if (f(ofrom(b + i)) < k) return NAN;
        if (b >= (uint64_t)i && f(ofrom(b - i)) >= k) return NAN;
    }
    return ofrom(b);
}

/* grid.c:28-39 latitude index as a function of q = z/Ri (libm asin). */
inline int td_of_q(double q, int ftd)
{
    double thi = asin(q);
    int td = (int)ceil((thi + M_PI / 2.0) / M_PI * ftd) - 1;
    if (td < 0) td = 0;
    if (td >= ftd) td = ftd - 1;
    return td;
}

/* Height index: ballistic.c:211-212 (uniform) or grid.c:91-99 (exponential). */
inline int ht_of_r(double Ri, double R, double unih, double alpha2, int exp_grid)
{
    int ht;
    if (exp_grid) {
        double arg = alpha2 * (Ri - R) / unih + 1.0;
        if (arg <= 1.0) return 0;
        ht = (int)ceil(log(arg) / log(1.0 + alpha2)) - 1;
    } else {
        ht = (int)ceil((Ri - R) / unih) - 1;
    }
    return ht < 0 ? 0 : ht;
}

struct HostTables {
    std::vector<double> tq, hr, cs, vt_speeds;
    std::vector<int32_t> vt_count;
    std::vector<double> cdf, enorm, fnorm;
    std::vector<double> face_sig;   /* sqrt(k_B T / mass) per finest face */
    std::vector<int32_t> fmap;      /* build_face_map, where a surface is uploaded */
    std::vector<int32_t> fmap_lat;  /* build_face_map's lattice (kept, so a reused HostTables reallocates nothing) */
    int level_off[12];
    int levels = 0, nfine = 0, nface_all = 0;
};

/* velocity.c:6-47 */
inline void velocity_table(double mmass, double mole, double k_B, double *speeds, int32_t *count)
{
    double dT = (240.0 - 20.0) / (DSMC_VT_NT - 1);
    double m = mmass / mole;
    for (int i = 0; i < DSMC_VT_NT; i++) {
        double T = 20.0 + i * dT;
        int n_v = 0;
        for (int j = 0; j <= 100; j++) {
            double vv = j * 10.0;
            double factor = 1.0 / (2.0 * M_PI) * m / (k_B * T);
            double aa = 4.0 * M_PI * vv * vv * pow(factor, 1.5) * exp(-m * 3.0 * vv * vv / (2.0 * k_B * T));
            aa *= 5e5;
            int cnt = (int)floor(aa);
            if (cnt == 0) continue;
            for (int kk = 0; kk < cnt && n_v < DSMC_VT_NV; kk++) speeds[i * DSMC_VT_NV + n_v++] = vv;
        }
        if (n_v == 0) n_v = 1;
        count[i] = n_v;
    }
}

/* Fills the scalar part of an environment and the host tables.  Returns 0 or a
 * dsmc_status.  `advanced` selects the compat advanced driver. */
inline int build_env(const dsmc_params *p, const dsmc_surface *s, int advanced, DsmcEnv *E,
                     HostTables *T)
{
    if (p->ftd <= 0 || p->ffd <= 0 || p->hd <= 0 || p->ftd > 4096 || p->ffd > 4096 || p->hd > 4096)
        return DSMC_ERR_ARG;
    if (!(p->dt > 0) || !(p->radius > 0) || !(p->max_height > p->radius) || !(p->unih > 0))
        return DSMC_ERR_ARG;
    if (p->exp_grid && !(p->alpha2 > 0)) return DSMC_ERR_ARG;
    memset(E, 0, sizeof(*E));
    E->mode = p->mode;
    E->advanced = advanced;
    E->ftd = p->ftd; E->ffd = p->ffd; E->hd = p->hd;
    E->exp_grid = advanced ? 1 : p->exp_grid;
    E->nparticles = p->nparticles;
    E->max_hop_steps = p->max_hop_steps > 0 ? p->max_hop_steps : INT64_MAX;
    E->max_steps = p->max_steps;
    E->seed = p->seed;
    E->stream = (uint32_t)p->stream;
    E->GM = p->G * p->M;
    E->R = p->radius;
    E->h = p->dt;
    double h2 = p->dt * p->dt;
    E->c1 = 1.5 * h2;
    E->c2 = h2 / 6.0;
    E->decay = exp(-p->dt / p->lifetime);
    E->dt_rot = p->period / p->ffd;
    E->ve = p->v_esc;
    E->max_height = p->max_height;
    E->w_min = p->w_min;
    E->t_freeze = p->t_freeze;
    E->alpha = p->alpha;
    E->period = p->period;
    E->mass = p->mmass / p->mole;
    E->k_B = p->k_B;
    E->lifetime = p->lifetime;
    dm_sincos(p->subsolar_lon, &E->sun_x, &E->sun_y);
    E->omega = 2.0 * M_PI / p->period;
    E->unih = p->unih;
    E->alpha2 = p->alpha2;
    E->sigma = p->cross_section;
    E->fit_deg = p->fit_degree;
    E->rec_cap_cell = p->record_cap_cell;
    E->rec_cap_total = p->record_cap_total;

    const int ftd = p->ftd, ffd = p->ffd, hd = p->hd;
    T->tq.assign(ftd + 1, 0.0);
    T->tq[0] = -INFINITY;
    T->tq[ftd] = INFINITY;
    for (int k = 1; k < ftd; k++) {
        double t = first_ge([ftd](double q) { return td_of_q(q, ftd); }, k, -1.0, 1.0);
        if (isnan(t)) return DSMC_ERR_ARG;
        T->tq[k] = t;
    }
    T->hr.assign(hd + 2, 0.0);
    T->hr[0] = -INFINITY;
    T->hr[hd + 1] = INFINITY;
    double R = p->radius, unih = p->unih, a2 = p->alpha2;
    int eg = E->exp_grid;
    double top = p->max_height * 4.0 + R;
    for (int k = 1; k <= hd; k++) {
        double t = first_ge([=](double r) { return ht_of_r(r, R, unih, a2, eg); }, k, R, top);
        if (isnan(t)) return DSMC_ERR_ARG;
        T->hr[k] = t;
    }
    T->cs.assign((size_t)(ffd + 1) * 2, 0.0);
    for (int k = 0; k <= ffd; k++) {
        double phi = 2.0 * M_PI * k / ffd;
        T->cs[2 * k] = cos(phi);
        T->cs[2 * k + 1] = sin(phi);
    }
    if (p->mode == DSMC_MODE_COMPAT) {
        if (!s->T_latlon || s->ftd != ftd || s->ffd != ffd) return DSMC_ERR_ARG;
        T->vt_speeds.assign(DSMC_VT_NT * DSMC_VT_NV, 0.0);
        T->vt_count.assign(DSMC_VT_NT, 0);
        velocity_table(p->mmass, p->mole, p->k_B, T->vt_speeds.data(), T->vt_count.data());
    } else {
        if (!s->vert || !s->faces || !s->face_T || !s->face_w || s->levels < 0 || s->levels > 10)
            return DSMC_ERR_ARG;
        int off = 0, nl = 20;
        for (int l = 0; l <= s->levels; l++) { T->level_off[l] = off; off += nl; nl *= 4; }
        T->levels = s->levels;
        T->nface_all = off;
        T->nfine = 20 << (2 * s->levels);
        T->cdf.assign(T->nfine, 0.0);
        double acc = 0.0;
        for (int i = 0; i < T->nfine; i++) {
            if (!(s->face_w[i] >= 0)) return DSMC_ERR_ARG;
            acc += s->face_w[i];
            T->cdf[i] = acc;
        }
        if (!(acc > 0)) return DSMC_ERR_ARG;
        /* the emission's thermal speed at each face, as the oracle computes it per
         * emission (correctly rounded division and square root on both sides) */
        T->face_sig.assign(T->nfine, 0.0);
        for (int i = 0; i < T->nfine; i++) T->face_sig[i] = sqrt(E->k_B * s->face_T[i] / E->mass);
        T->enorm.assign((size_t)off * 9, 0.0);
        for (int f = 0; f < off; f++) {
            const int32_t *fv = s->faces + 3 * (size_t)f;
            for (int j = 0; j < 3; j++) if (fv[j] < 0 || fv[j] >= s->nvert) return DSMC_ERR_ARG;
            const double *v[3] = {s->vert + 3 * (size_t)fv[0], s->vert + 3 * (size_t)fv[1],
                                  s->vert + 3 * (size_t)fv[2]};
            for (int e = 0; e < 3; e++) {
                const double *a = v[e], *b = v[(e + 1) % 3];
                double *n = &T->enorm[(size_t)f * 9 + 3 * e];
                n[0] = a[1] * b[2] - a[2] * b[1];
                n[1] = a[2] * b[0] - a[0] * b[2];
                n[2] = a[0] * b[1] - a[1] * b[0];
            }
        }
        /* unit normals of the finest faces (the locate_face_near guess test) */
        T->fnorm.assign((size_t)T->nfine * 9, 0.0);
        const size_t f0 = (size_t)T->level_off[T->levels];
        for (size_t f = 0; f < (size_t)T->nfine; f++)
            for (int e = 0; e < 3; e++) {
                const double *n = &T->enorm[(f0 + f) * 9 + 3 * e];
                double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (!(len > 0)) return DSMC_ERR_ARG;
                for (int j = 0; j < 3; j++) T->fnorm[f * 9 + 3 * e + j] = n[j] / len;
            }
        for (int l = 0; l <= T->levels; l++) E->level_off[l] = T->level_off[l];
        E->levels = T->levels;
        E->nfine = T->nfine;
    }
    return DSMC_OK;
}

/* Cube-map resolution of the face map: about 1.25 finest faces per cell (2-3
 * candidates in a typical cell, at most 6 around a vertex), 8 <= G <= 128. */
inline int face_map_grid(int nfine)
{
    int g = (int)lround(sqrt(nfine / 1.25));
    return g < 8 ? 8 : (g > 128 ? 128 : g);
}

/* The face map of a physical surface (fmap_cell's cells, DSMC_FMAP_K candidates
 * per cell, -1 padded): the finest faces holding the points of a lattice of
 * 4 x 4 sub-cells per cell (corners and edges included, so neighbouring cells
 * share their boundary points), most frequent first, lowest index on ties.  A
 * face the lattice misses, or a cell with more than DSMC_FMAP_K faces, costs
 * only the full search: locate_face_near tests every candidate exactly.
 * `E` holds the surface scalars; enorm / fnorm are read from T. */
inline void build_face_map(const DsmcEnv &E, const HostTables &T, int G, std::vector<int32_t> &out,
                           std::vector<int32_t> *lattice = nullptr)
{
    const int S = 4, L = G * S + 1;
    DsmcEnv H = E;
    H.enorm = T.enorm.data();
    H.fnorm = T.fnorm.data();
    H.fmap = nullptr;
    std::vector<int32_t> own;
    std::vector<int32_t> &lat = lattice ? *lattice : own;
    lat.assign((size_t)L * L, 0);
    out.assign((size_t)6 * G * G * DSMC_FMAP_K, -1);
    for (int cf = 0; cf < 6; cf++) {
        const double sg = (cf & 1) ? -1.0 : 1.0;
        const int axis = cf >> 1;
        int hint = -1;
        for (int j = 0; j < L; j++)
            for (int i = 0; i < L; i++) {
                const double u = -1.0 + 2.0 * i / (L - 1), v = -1.0 + 2.0 * j / (L - 1);
                double d[3];
                if (axis == 0) { d[0] = sg; d[1] = u; d[2] = v; }
                else if (axis == 1) { d[0] = v; d[1] = sg; d[2] = u; }
                else { d[0] = u; d[1] = v; d[2] = sg; }
                const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                hint = locate_face_near(H, hint, d[0], d[1], d[2], len);
                lat[(size_t)j * L + i] = hint;
            }
        for (int cj = 0; cj < G; cj++)
            for (int ci = 0; ci < G; ci++) {
                int face[(S + 1) * (S + 1)], cnt[(S + 1) * (S + 1)], n = 0;
                for (int j = cj * S; j <= (cj + 1) * S; j++)
                    for (int i = ci * S; i <= (ci + 1) * S; i++) {
                        const int f = lat[(size_t)j * L + i];
                        int k = 0;
                        while (k < n && face[k] != f) k++;
                        if (k == n) { face[n] = f; cnt[n++] = 0; }
                        cnt[k]++;
                    }
                int32_t *o = &out[((size_t)(cf * G + cj) * G + ci) * DSMC_FMAP_K];
                for (int k = 0; k < DSMC_FMAP_K && k < n; k++) {
                    int b = -1;
                    for (int q = 0; q < n; q++)
                        if (cnt[q] > 0 && (b < 0 || cnt[q] > cnt[b] || (cnt[q] == cnt[b] && face[q] < face[b]))) b = q;
                    o[k] = face[b];
                    cnt[b] = 0;
                }
            }
    }
}

/* fnorm's rows of the first four candidates of every face-map cell (a -1
 * candidate takes the first candidate's row; a cell without candidates, zeros),
 * [cells][4][9]: the layout of DsmcEnv::fmap_n. */
inline void build_face_map_normals(const std::vector<int32_t> &fmap, const std::vector<double> &fnorm, int G,
                                   std::vector<double> &out)
{
    const size_t cells = (size_t)6 * G * G;
    out.assign(cells * 36, 0.0);
    for (size_t c = 0; c < cells; c++) {
        const int32_t *f = &fmap[c * DSMC_FMAP_K];
        if (f[0] < 0) continue;
        for (int j = 0; j < 4; j++) {
            const int32_t face = f[j] >= 0 ? f[j] : f[0];
            for (int k = 0; k < 9; k++) out[c * 36 + 9 * j + k] = fnorm[(size_t)face * 9 + k];
        }
    }
}

}  // namespace dsmc

#endif /* DSMC_TABLES_H */
