/*
 * shadow.hip - dsmc_suite's Monte-Carlo shadow flux (src/raytracer/
 * shadow_maker.c: mesh_load, mesh_compute_normals, random_point_on_triangle,
 * flux_compute_embree, run_raytracer) on the device.
 *
 * One wavefront per (face, time step): the sun direction of the step, the face's
 * cosine factor, then the Nmonte points the reference draws on the face with
 * glibc rand_r (seed t*nface + nf + 42), one per lane (each lane jumps to its
 * point's place in the sequence: rand_r is three steps of a linear
 * congruential generator, and k steps of it are one affine map), an any-hit
 * occlusion ray toward the sun from each, and the wavefront's count of hits.  The reference queries Embree's single-
 * precision BVH; here the test is double-precision Moller-Trumbore (closed
 * triangles, hit when t > 0) and the acceleration structure a binary BVH built
 * on the host (median split of face centroids along the longest axis, at most
 * four faces per leaf, boxes padded by 1e-9 of the mesh extent so the box test
 * never rejects a face the triangle test would accept: the answer is that of a
 * test against every face).  Normals, sample points, sun directions and the
 * flux follow shadow_maker.c operation for operation in double precision
 * (sin / cos from dsmc_math.h), so the result equals the oracle's device-math
 * build bit for bit.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <vector>

#include "../../include/dsmc_hip.h"
#include "dsmc_math.h"
#include "dsmc_host.h"

#define HIP_OK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "dsmc_hip: %s failed: %s (%s:%d)\n", #expr,                  \
                    hipGetErrorString(e_), __FILE__, __LINE__);                           \
            rc = DSMC_ERR_DEVICE;                                                         \
            goto done;                                                                    \
        }                                                                                 \
    } while (0)

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { BVH_LEAF = 4, BVH_STACK = 64 };

struct BvhNode {
    double lo[3], hi[3];
    int left, right;        /* children (inner node) */
    int first, count;       /* faces [first, first + count) of the permuted order (leaf: count > 0) */
};

/* glibc's rand_r (stdlib/rand_r.c): three steps of the LCG 1103515245 n + 12345,
 * 11 + 10 + 10 bits of their high halves. */
__host__ __device__ static inline int rand_r_glibc(unsigned int *seed)
{
    unsigned int next = *seed;
    int result;
    next *= 1103515245u;
    next += 12345u;
    result = (unsigned int)(next / 65536) % 2048;
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (unsigned int)(next / 65536) % 1024;
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (unsigned int)(next / 65536) % 1024;
    *seed = next;
    return result;
}

/* The generator state k steps of x -> 1103515245 x + 12345 (mod 2^32) after s:
 * powers of one affine map, composed by repeated squaring. */
__host__ __device__ static inline unsigned int lcg_skip(unsigned int s, unsigned int k)
{
    unsigned int A = 1u, C = 0u, a = 1103515245u, c = 12345u;
    while (k) {
        if (k & 1u) { A = a * A; C = a * C + c; }
        c = a * c + c;
        a = a * a;
        k >>= 1;
    }
    return A * s + C;
}

/* n values of the sequence: in turn (dsmc_rand_r_sequence), or each reached by
 * jumping 3 i generator steps from `seed`, the kernel's way (dsmc_rand_r_at). */
extern "C" int dsmc_rand_r_sequence(unsigned int seed, int n, int *out)
{
    if (n < 0 || (n && !out)) return DSMC_ERR_ARG;
    for (int i = 0; i < n; i++) out[i] = rand_r_glibc(&seed);
    return DSMC_OK;
}

extern "C" int dsmc_rand_r_at(unsigned int seed, int n, int *out)
{
    if (n < 0 || (n && !out)) return DSMC_ERR_ARG;
    for (int i = 0; i < n; i++) {
        unsigned int s = lcg_skip(seed, 3u * (unsigned int)i);
        out[i] = rand_r_glibc(&s);
    }
    return DSMC_OK;
}

/* Moller-Trumbore: the ray o + t d (t > 0) meets the closed triangle v0, v0 +
 * e1, v0 + e2. */
__device__ static inline bool tri_hit(const double o[3], const double d[3], const double *v)
{
    const double *v0 = v, *e1 = v + 3, *e2 = v + 6;
    const double px = d[1] * e2[2] - d[2] * e2[1], py = d[2] * e2[0] - d[0] * e2[2], pz = d[0] * e2[1] - d[1] * e2[0];
    const double det = e1[0] * px + e1[1] * py + e1[2] * pz;
    if (det == 0.0) return false;
    const double sx = o[0] - v0[0], sy = o[1] - v0[1], sz = o[2] - v0[2];
    const double u = (sx * px + sy * py + sz * pz) / det;
    if (!(u >= 0.0 && u <= 1.0)) return false;
    const double qx = sy * e1[2] - sz * e1[1], qy = sz * e1[0] - sx * e1[2], qz = sx * e1[1] - sy * e1[0];
    const double v_ = (d[0] * qx + d[1] * qy + d[2] * qz) / det;
    if (!(v_ >= 0.0 && u + v_ <= 1.0)) return false;
    const double t = (e2[0] * qx + e2[1] * qy + e2[2] * qz) / det;
    return t > 0.0;
}

/* The ray meets the box at some t >= 0 (id = 1 / d; the boxes are padded, so
 * the rounding of the products cannot reject a face the triangle test accepts). */
__device__ static inline bool box_hit(const double o[3], const double d[3], const double id[3], const BvhNode &n)
{
    double t0 = 0.0, t1 = INFINITY;
    for (int k = 0; k < 3; k++) {
        if (d[k] == 0.0) {
            if (o[k] < n.lo[k] || o[k] > n.hi[k]) return false;
            continue;
        }
        double a = (n.lo[k] - o[k]) * id[k], b = (n.hi[k] - o[k]) * id[k];
        if (a > b) { double s = a; a = b; b = s; }
        t0 = a > t0 ? a : t0;
        t1 = b < t1 ? b : t1;
        if (t1 < t0) return false;
    }
    return true;
}

__device__ static bool occluded(const double o[3], const double d[3], const double id[3], const BvhNode *nodes,
                                const double *tris)
{
    int stack[BVH_STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const BvhNode n = nodes[stack[--sp]];
        if (!box_hit(o, d, id, n)) continue;
        if (n.count > 0) {
            for (int i = 0; i < n.count; i++)
                if (tri_hit(o, d, tris + (size_t)(n.first + i) * 9)) return true;
        } else {
            stack[sp++] = n.left;        /* depth <= BVH_STACK - 1, checked at build */
            stack[sp++] = n.right;
        }
    }
    return false;
}

/* face: [nface][12] = a, b, c, normal; F[nf * tt + t] (shadow_maker.c:198-239).
 * One wavefront per (face, step); blocks of whole wavefronts. */
__global__ void __launch_bounds__(256) shadow_kernel(int nface, int tt, int nmonte, double za, const double *face,
                                                     const BvhNode *nodes, const double *tris, double *F)
{
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;     /* uniform in the wavefront */
    const int lane = (int)(threadIdx.x & 63);
    if (i >= (int64_t)nface * tt) return;
    const int nf = (int)(i / tt), t = (int)(i % tt);
    const double fi = ((t + 0.5) / tt) * 2.0 * M_PI + M_PI;
    const double cz = dm_cos(-za);
    const double sun[3] = {cz * dm_sin(fi), cz * dm_cos(fi), dm_sin(-za)};
    const double isun[3] = {1.0 / sun[0], 1.0 / sun[1], 1.0 / sun[2]};
    const double *f = face + (size_t)nf * 12;
    const double c6 = sun[0] * f[9] + sun[1] * f[10] + sun[2] * f[11];
    if (c6 < 0) {
        if (lane == 0) F[i] = 0.0;
        return;
    }
    const unsigned int seed0 = (unsigned int)(t * nface + nf + 42);
    const double eps = (double)1e-3f;
    int shadow = 0;
    for (int base = 0; base < nmonte; base += 64) {
        const int mc = base + lane;
        bool hit = false;
        if (mc < nmonte) {
            unsigned int seed = lcg_skip(seed0, 6u * (unsigned int)mc);   /* two rand_r calls per point */
            const double r1 = sqrt((double)rand_r_glibc(&seed) / 2147483647);
            const double r2 = (double)rand_r_glibc(&seed) / 2147483647;
            const double wa = 1.0 - r1, wb = r1 * (1.0 - r2), wc = r1 * r2;
            double o[3];
            for (int k = 0; k < 3; k++) o[k] = f[k] * wa + f[3 + k] * wb + f[6 + k] * wc + eps * f[9 + k];
            hit = occluded(o, sun, isun, nodes, tris);
        }
        shadow += __popcll(__ballot(hit));
    }
    if (lane == 0) F[i] = (1.0 - (double)shadow / nmonte) * c6;
}

/* ------------------------------------------------------------------ host */
/* shadow_maker.c:70-110 (vec3_normalize multiplies by 1 / length) */
static void normalize3(const double v[3], double out[3])
{
    const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l < 1e-30) { out[0] = out[1] = out[2] = 0.0; return; }
    const double s = 1.0 / l;
    out[0] = v[0] * s; out[1] = v[1] * s; out[2] = v[2] * s;
}

static void face_record(const double *vert, const int32_t *fc, double rec[12])
{
    const double *a = vert + 3 * (size_t)fc[0], *b = vert + 3 * (size_t)fc[1], *c = vert + 3 * (size_t)fc[2];
    double cen[3], c1[3], c2[3], c1n[3], c2n[3], c3[3], nrm[3];
    for (int k = 0; k < 3; k++) {
        cen[k] = (a[k] + b[k] + c[k]) / 3.0;
        c1[k] = b[k] - a[k];
        c2[k] = c[k] - a[k];
    }
    normalize3(c2, c2n);
    normalize3(c1, c1n);
    c3[0] = c2n[1] * c1n[2] - c2n[2] * c1n[1];
    c3[1] = c2n[2] * c1n[0] - c2n[0] * c1n[2];
    c3[2] = c2n[0] * c1n[1] - c2n[1] * c1n[0];
    normalize3(c3, nrm);
    if (nrm[0] * cen[0] + nrm[1] * cen[1] + nrm[2] * cen[2] < 0)
        for (int k = 0; k < 3; k++) nrm[k] = nrm[k] * -1.0;
    for (int k = 0; k < 3; k++) { rec[k] = a[k]; rec[3 + k] = b[k]; rec[6 + k] = c[k]; rec[9 + k] = nrm[k]; }
}

struct BvhBuild {
    const double *vert;
    const int32_t *faces;
    std::vector<int> order;
    std::vector<double> cen;       /* [nface][3] */
    std::vector<BvhNode> nodes;
    double pad = 0.0;
    int depth = 0;

    int build(int first, int count, int level)
    {
        depth = std::max(depth, level);
        const int id = (int)nodes.size();
        nodes.push_back(BvhNode());
        BvhNode n;
        for (int k = 0; k < 3; k++) { n.lo[k] = INFINITY; n.hi[k] = -INFINITY; }
        double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = first; i < first + count; i++) {
            const int f = order[i];
            for (int j = 0; j < 3; j++) {
                const double *v = vert + 3 * (size_t)faces[3 * (size_t)f + j];
                for (int k = 0; k < 3; k++) { n.lo[k] = std::min(n.lo[k], v[k]); n.hi[k] = std::max(n.hi[k], v[k]); }
            }
            for (int k = 0; k < 3; k++) {
                clo[k] = std::min(clo[k], cen[3 * (size_t)f + k]);
                chi[k] = std::max(chi[k], cen[3 * (size_t)f + k]);
            }
        }
        for (int k = 0; k < 3; k++) { n.lo[k] -= pad; n.hi[k] += pad; }
        if (count <= BVH_LEAF) {
            n.left = n.right = -1;
            n.first = first;
            n.count = count;
        } else {
            int ax = 0;
            for (int k = 1; k < 3; k++) if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
            const int mid = first + count / 2;
            std::nth_element(order.begin() + first, order.begin() + mid, order.begin() + first + count,
                             [&](int p, int q) {
                                 const double a = cen[3 * (size_t)p + ax], b = cen[3 * (size_t)q + ax];
                                 return a < b || (a == b && p < q);
                             });
            n.first = 0;
            n.count = 0;
            n.left = build(first, mid - first, level + 1);
            n.right = build(mid, first + count - mid, level + 1);
        }
        nodes[id] = n;
        return id;
    }
};

static int shadow_run(int npoint, const double *vert, int nface, const int32_t *faces, double za_deg, int tt,
                      int nmonte, double *F)
{
    if (npoint <= 0 || nface <= 0 || !vert || !faces || tt <= 0 || nmonte <= 0 || !F) return DSMC_ERR_ARG;
    if ((int64_t)nface * tt > (int64_t)INT32_MAX) return DSMC_ERR_ARG;      /* one lane per (face, step) */
    for (int64_t i = 0; i < 3 * (int64_t)nface; i++)
        if (faces[i] < 0 || faces[i] >= npoint) return DSMC_ERR_ARG;
    int rc = dsmc::ensure_device();
    if (rc) return rc;
    const double za = za_deg / 180.0 * M_PI;
    std::vector<double> face((size_t)nface * 12);
    for (int f = 0; f < nface; f++) face_record(vert, faces + 3 * (size_t)f, &face[(size_t)f * 12]);
    BvhBuild B;
    B.vert = vert;
    B.faces = faces;
    B.order.resize(nface);
    B.cen.resize((size_t)nface * 3);
    double ext = 0.0, lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < npoint; i++)
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], vert[3 * i + k]); hi[k] = std::max(hi[k], vert[3 * i + k]); }
    for (int k = 0; k < 3; k++) ext = std::max(ext, hi[k] - lo[k]);
    if (!(ext < INFINITY)) return DSMC_ERR_ARG;
    B.pad = 1e-9 * ext + 1e-300;
    for (int f = 0; f < nface; f++) {
        B.order[f] = f;
        for (int k = 0; k < 3; k++) B.cen[3 * (size_t)f + k] = face[(size_t)f * 12 + k] / 3.0 +
                                                               face[(size_t)f * 12 + 3 + k] / 3.0 +
                                                               face[(size_t)f * 12 + 6 + k] / 3.0;
    }
    B.nodes.reserve(2 * (size_t)nface / BVH_LEAF + 2);
    B.build(0, nface, 1);
    if (B.depth > BVH_STACK - 2) return DSMC_ERR_ARG;            /* the traversal stack holds depth + 1 */
    std::vector<double> tris((size_t)nface * 9);
    for (int i = 0; i < nface; i++) {
        const double *r = &face[(size_t)B.order[i] * 12];
        double *t = &tris[(size_t)i * 9];
        for (int k = 0; k < 3; k++) { t[k] = r[k]; t[3 + k] = r[3 + k] - r[k]; t[6 + k] = r[6 + k] - r[k]; }
    }
    double *d_face = nullptr, *d_tris = nullptr, *d_F = nullptr;
    BvhNode *d_nodes = nullptr;
    const size_t nF = (size_t)nface * tt;
    HIP_OK(hipMalloc((void **)&d_face, face.size() * 8));
    HIP_OK(hipMalloc((void **)&d_tris, tris.size() * 8));
    HIP_OK(hipMalloc((void **)&d_nodes, B.nodes.size() * sizeof(BvhNode)));
    HIP_OK(hipMalloc((void **)&d_F, nF * 8));
    HIP_OK(hipMemcpy(d_face, face.data(), face.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_tris, tris.data(), tris.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_nodes, B.nodes.data(), B.nodes.size() * sizeof(BvhNode), hipMemcpyHostToDevice));
    /* four (face, step) wavefronts per 256-thread block */
    hipLaunchKernelGGL(shadow_kernel, dim3((unsigned)((nF + 3) / 4)), dim3(256), 0, 0, nface, tt, nmonte, za, d_face,
                       d_nodes, d_tris, d_F);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(F, d_F, nF * 8, hipMemcpyDeviceToHost));
done:
    for (void *p : {(void *)d_face, (void *)d_tris, (void *)d_nodes, (void *)d_F})
        if (p) (void)hipFree(p);
    return rc;
}

extern "C" int dsmc_shadow_flux(int npoint, const double *vert, int nface, const int32_t *faces, double za_deg, int tt,
                                int nmonte, double *F)
{
    return shadow_run(npoint, vert, nface, faces, za_deg, tt, nmonte, F);
}

/* shadow_maker.c:23-55 (vertex lines "x y z", face lines of 1-based indices)
 * and :242-271 */
extern "C" void run_raytracer(const char *mesh_vertex_file, const char *mesh_face_file, int npoint, int nface,
                              double za_deg, int tt, int Nmonte, const char *output_file, int nthreads)
{
    (void)nthreads;
    printf("Loading mesh: %d vertices, %d faces\n", npoint, nface);
    if (npoint <= 0 || nface <= 0) {
        fprintf(stderr, "run_raytracer: invalid mesh sizes\n");
        return;
    }
    std::vector<double> vert((size_t)npoint * 3);
    std::vector<int32_t> faces((size_t)nface * 3);
    FILE *fp = fopen(mesh_vertex_file, "r");
    if (!fp) { fprintf(stderr, "Cannot open %s\n", mesh_vertex_file); return; }
    for (int i = 0; i < npoint; i++)
        if (fscanf(fp, "%lf %lf %lf", &vert[3 * i], &vert[3 * i + 1], &vert[3 * i + 2]) != 3) {
            fprintf(stderr, "%s: short vertex file\n", mesh_vertex_file);
            fclose(fp);
            return;
        }
    fclose(fp);
    fp = fopen(mesh_face_file, "r");
    if (!fp) { fprintf(stderr, "Cannot open %s\n", mesh_face_file); return; }
    for (int i = 0; i < nface; i++) {
        int a, b, c;
        if (fscanf(fp, "%d %d %d", &a, &b, &c) != 3) {
            fprintf(stderr, "%s: short face file\n", mesh_face_file);
            fclose(fp);
            return;
        }
        faces[3 * i] = a - 1; faces[3 * i + 1] = b - 1; faces[3 * i + 2] = c - 1;
    }
    fclose(fp);
    printf("Building BVH...\n");
    printf("BVH built.\n");
    printf("Computing flux with %d MC samples (device)...\n", Nmonte);
    std::vector<double> F((size_t)nface * tt);
    int rc = shadow_run(npoint, vert.data(), nface, faces.data(), za_deg, tt, Nmonte, F.data());
    if (rc) {
        fprintf(stderr, "run_raytracer: failed (%d)\n", rc);
        return;
    }
    write_flux_dat(output_file, F.data(), nface, tt);
    printf("Flux written to %s\n", output_file);
}
