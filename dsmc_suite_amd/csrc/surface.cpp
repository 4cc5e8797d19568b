/*
 * surface.cpp - the icosphere surface of the benchmark configurations.
 *
 * Subdivision follows the algorithm of scripts/vesta_raytrace_gif.py:35-102: the
 * 12 normalised icosahedron vertices and 20 faces, then per level every face
 * (a, b, c) is split at its normalised edge midpoints into [a,ab,ca], [b,bc,ab],
 * [c,ca,bc], [ab,bc,ca], midpoints numbered in order of first use.  Children of
 * face f are therefore faces 4f..4f+3 of the next level, the hierarchy the device
 * uses to locate impact points.
 *
 * Face temperature: T = t_night + t_amp * max(cos(sza), 0)^0.25 at the normalised
 * face centroid, sun in the equatorial plane at longitude subsolar_lon (the
 * longitude convention of ballistic.c:114-117: s = (cos thi sin fi, cos thi cos
 * fi, sin thi)).  Launch weight: flat face area x Hertz-Knudsen sublimation flux
 * P(T) / sqrt(2 pi m k T) with the Antoine vapour pressure of
 * src/thermal/thermal_model.c:50-55 and its H2O coefficients (:33-38).
 */
#include <math.h>
#include <string.h>
#include <map>
#include <vector>

#include "../../include/dsmc_hip.h"
#include "dsmc_tables.h"

extern "C" int dsmc_icosphere_sizes(int levels, int *nvert, int *nface_all, int *nface_fine)
{
    if (levels < 0 || levels > 10) return DSMC_ERR_ARG;
    int nf = 20, all = 0;
    for (int l = 0; l <= levels; l++) { all += nf; if (l < levels) nf *= 4; }
    if (nvert) *nvert = nf / 2 + 2;
    if (nface_all) *nface_all = all;
    if (nface_fine) *nface_fine = nf;
    return DSMC_OK;
}

static double antoine_pressure(double T)
{
    const double p0 = -2445.5646, p1 = -6.757169, p2 = -0.01677006, p3 = 1.20514e-5, p4 = 8.2312;
    double lg = p0 / T + p1 + p2 * T + p3 * T * T + p4 * log10(T);
    return 133.32237 * pow(10.0, lg);
}

extern "C" int dsmc_icosphere_build(int levels, double subsolar_lon, double t_night, double t_amp,
                                    double molecule_mass, double k_B, double *vert, int32_t *faces,
                                    double *face_T, double *face_w)
{
    int nv, nall, nfine;
    if (dsmc_icosphere_sizes(levels, &nv, &nall, &nfine) || !vert || !faces) return DSMC_ERR_ARG;
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int f0[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                           {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                           {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    std::vector<double> V;
    V.reserve((size_t)nv * 3);
    for (int i = 0; i < 12; i++) {
        double n = sqrt(base[i][0] * base[i][0] + base[i][1] * base[i][1] + base[i][2] * base[i][2]);
        V.push_back(base[i][0] / n); V.push_back(base[i][1] / n); V.push_back(base[i][2] / n);
    }
    std::vector<int32_t> F;
    for (int i = 0; i < 20; i++) for (int j = 0; j < 3; j++) F.push_back(f0[i][j]);
    size_t level_start = 0;
    for (int l = 0; l < levels; l++) {
        std::map<std::pair<int, int>, int> mid;
        size_t nprev = (F.size() - level_start) / 3;
        size_t start = F.size();
        auto get_mid = [&](int a, int b) {
            std::pair<int, int> key(a < b ? a : b, a < b ? b : a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double m[3];
            for (int k = 0; k < 3; k++) m[k] = (V[3 * a + k] + V[3 * b + k]) / 2.0;
            double n = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            int idx = (int)(V.size() / 3);
            for (int k = 0; k < 3; k++) V.push_back(m[k] / n);
            mid[key] = idx;
            return idx;
        };
        for (size_t f = 0; f < nprev; f++) {
            int a = F[level_start + 3 * f], b = F[level_start + 3 * f + 1], c = F[level_start + 3 * f + 2];
            int ab = get_mid(a, b), bc = get_mid(b, c), ca = get_mid(c, a);
            int kids[4][3] = {{a, ab, ca}, {b, bc, ab}, {c, ca, bc}, {ab, bc, ca}};
            for (int k = 0; k < 4; k++) for (int j = 0; j < 3; j++) F.push_back(kids[k][j]);
        }
        level_start = start;
    }
    if ((int)(V.size() / 3) != nv || (int)(F.size() / 3) != nall) return DSMC_ERR_ARG;
    memcpy(vert, V.data(), V.size() * sizeof(double));
    memcpy(faces, F.data(), F.size() * sizeof(int32_t));
    double sx = sin(subsolar_lon), sy = cos(subsolar_lon);
    const int32_t *fine = faces + 3 * (size_t)(nall - nfine);
    for (int f = 0; f < nfine; f++) {
        const double *a = vert + 3 * fine[3 * f], *b = vert + 3 * fine[3 * f + 1], *c = vert + 3 * fine[3 * f + 2];
        double cx = a[0] + b[0] + c[0], cy = a[1] + b[1] + c[1], cz = a[2] + b[2] + c[2];
        double cn = sqrt(cx * cx + cy * cy + cz * cz);
        double mu = (cx * sx + cy * sy) / cn;
        double T = t_night + t_amp * pow(mu > 0.0 ? mu : 0.0, 0.25);
        if (face_T) face_T[f] = T;
        if (face_w) {
            double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                            e1[0] * e2[1] - e1[1] * e2[0]};
            double area = 0.5 * sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            double flux = antoine_pressure(T) / sqrt(2.0 * M_PI * molecule_mass * k_B * T);
            face_w[f] = area * flux;
        }
    }
    return DSMC_OK;
}

extern "C" int dsmc_icosphere_rasterise(const dsmc_surface *s, int ftd, int ffd, double *T_latlon)
{
    if (!s || !s->vert || !s->faces || !s->face_T || ftd <= 0 || ffd <= 0 || !T_latlon || s->levels < 0 ||
        s->levels > 10)
        return DSMC_ERR_ARG;
    dsmc_params p;
    dsmc_params_preset(&p, DSMC_PRESET_CERES_PHYSICAL, ftd, ffd);
    std::vector<double> w(20 << (2 * s->levels), 1.0);
    dsmc_surface s2 = *s;
    s2.face_w = w.data();
    DsmcEnv E;
    dsmc::HostTables T;
    int rc = dsmc::build_env(&p, &s2, 0, &E, &T);
    if (rc) return rc;
    E.enorm = T.enorm.data();
    for (int td = 0; td < ftd; td++) {
        double thi = (td + 0.5) / ftd * M_PI - M_PI / 2.0;
        for (int fd = 0; fd < ffd; fd++) {
            double fi = (fd + 0.5) / ffd * 2.0 * M_PI;
            double d[3] = {cos(thi) * sin(fi), cos(thi) * cos(fi), sin(thi)};
            int f = locate_face(E, d[0], d[1], d[2]);
            T_latlon[td * ffd + fd] = s->face_T[f];
        }
    }
    return DSMC_OK;
}
