// Scene generators for the benchmark workloads.
#include <cmath>
#include <cstring>
#include <vector>

#include "spg_internal.h"

using namespace spg;

namespace {

// CreateUnitVertex, src/mesh_generation.cpp:209-216: position normalized,
// normal = position, uv = (x, y).
void unit_vertex(float *v8, float x, float y, float z)
{
    v3 p = normalize(v3{x, y, z});
    v8[0] = p.x; v8[1] = p.y; v8[2] = p.z;
    v8[3] = p.x; v8[4] = p.y; v8[5] = p.z;
    v8[6] = p.x; v8[7] = p.y;
}

// splitmix64 (Steele, Lea, Flood 2014) for reproducible scene seeds
struct SplitMix
{
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

struct Sky
{
    double sun[3];
    double cosSun;
    float sunRadiance;
    void eval(const double *d, float *rgb) const
    {
        double c = d[0] * sun[0] + d[1] * sun[1] + d[2] * sun[2];
        if (c >= cosSun)
        {
            rgb[0] = rgb[1] = rgb[2] = sunRadiance;
            return;
        }
        if (d[1] >= 0.0)
        {
            double t = d[1];
            const double horizon[3] = {0.9, 1.0, 1.1}, zenith[3] = {0.25, 0.45, 0.9};
            for (int k = 0; k < 3; ++k)
                rgb[k] = (float)(horizon[k] * (1.0 - t) + zenith[k] * t);
        }
        else
        {
            rgb[0] = 0.25f; rgb[1] = 0.22f; rgb[2] = 0.2f;
        }
    }
};

bool make_sky(const float sunDir[3], float angR, float radiance, Sky &s)
{
    double l = sqrt((double)sunDir[0] * sunDir[0] + (double)sunDir[1] * sunDir[1] +
                    (double)sunDir[2] * sunDir[2]);
    if (!(l > 0.0)) return false;
    for (int k = 0; k < 3; ++k) s.sun[k] = sunDir[k] / l;
    s.cosSun = cos((double)angR);
    s.sunRadiance = radiance;
    return true;
}

}  // namespace

extern "C" {

// CreateIcosahedronMesh, src/mesh_generation.cpp:220-337: the 12 vertices of
// an icosahedron normalized to the unit sphere and its 20 faces; each
// subdivision level adds three new (unshared) midpoint vertices per face and
// replaces it by (i,m0,m2) (j,m1,m0) (k,m2,m1) (m0,m1,m2) where m0, m1, m2
// are the midpoints of edges ij, jk, ki.
int spg_gen_icosphere(uint32_t level, float *verts8, uint32_t vertCap,
                      uint32_t *idx, uint32_t idxCap, uint32_t *nv,
                      uint32_t *ni)
{
    if (!verts8 || !idx || !nv || !ni)
        return fail(SPG_E_INVALID, "spg_gen_icosphere: NULL argument");
    uint64_t faces = 20;
    uint64_t verts = 12;
    for (uint32_t l = 0; l < level; ++l)
    {
        verts += 3 * faces;
        faces *= 4;
    }
    if (verts > vertCap || 3 * faces > idxCap || level > 10)
        return fail(SPG_E_CAPACITY, "spg_gen_icosphere: buffers too small");
    const float t = (1.0f + sqrtf(5.0f)) * 0.5f;
    const float base[12][3] = {{-1, t, 0}, {1, t, 0},  {-1, -t, 0}, {1, -t, 0},
                               {0, -1, t}, {0, 1, t},  {0, -1, -t}, {0, 1, -t},
                               {t, 0, -1}, {t, 0, 1},  {-t, 0, -1}, {-t, 0, 1}};
    static const uint32_t F[20][3] = {
        {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
        {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
        {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
        {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    uint32_t vc = 0;
    for (int i = 0; i < 12; ++i)
        unit_vertex(verts8 + 8 * vc++, base[i][0], base[i][1], base[i][2]);
    std::vector<uint32_t> cur(&F[0][0], &F[0][0] + 60), nxt;
    for (uint32_t l = 0; l < level; ++l)
    {
        nxt.clear();
        nxt.reserve(cur.size() * 4);
        for (size_t f = 0; f < cur.size(); f += 3)
        {
            uint32_t fi = cur[f], fj = cur[f + 1], fk = cur[f + 2];
            uint32_t e[6] = {fi, fj, fj, fk, fk, fi};
            uint32_t m[3];
            for (int k = 0; k < 3; ++k)
            {
                const float *a = verts8 + 8 * e[2 * k];
                const float *b = verts8 + 8 * e[2 * k + 1];
                unit_vertex(verts8 + 8 * vc, a[0] + b[0], a[1] + b[1], a[2] + b[2]);
                m[k] = vc++;
            }
            const uint32_t q[12] = {fi, m[0], m[2], fj, m[1], m[0],
                                    fk, m[2], m[1], m[0], m[1], m[2]};
            nxt.insert(nxt.end(), q, q + 12);
        }
        cur.swap(nxt);
    }
    memcpy(idx, cur.data(), 4 * cur.size());
    *nv = vc;
    *ni = (uint32_t)cur.size();
    return SPG_OK;
}

// CreatePlaneMesh, src/mesh_generation.cpp:1-21: unit quad in XY, normal +Z.
int spg_gen_plane(float *verts8, uint32_t *idx)
{
    if (!verts8 || !idx) return fail(SPG_E_INVALID, "spg_gen_plane: NULL");
    const float V[32] = {-0.5f, -0.5f, 0, 0, 0, 1, 0, 0, 0.5f, -0.5f, 0, 0, 0, 1, 1, 0,
                         0.5f,  0.5f,  0, 0, 0, 1, 1, 1, -0.5f, 0.5f, 0, 0, 0, 1, 0, 1};
    const uint32_t I[6] = {0, 1, 2, 2, 3, 0};
    memcpy(verts8, V, sizeof V);
    memcpy(idx, I, sizeof I);
    return SPG_OK;
}

int spg_gen_triangle_soup(uint64_t seed, uint32_t n, float mu, float sigma,
                          float *tris9)
{
    if (!tris9 && n) return fail(SPG_E_INVALID, "spg_gen_triangle_soup: NULL");
    SplitMix rng{seed};
    for (uint32_t i = 0; i < n; ++i)
    {
        double c[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
        double L = exp((double)mu + (double)sigma * rng.normal());
        // random unit normal (uniform on the sphere) and in-plane angle
        double z = 2.0 * rng.uniform() - 1.0, ph = 6.283185307179586 * rng.uniform();
        double r = sqrt(std::max(0.0, 1.0 - z * z));
        double nrm[3] = {r * cos(ph), r * sin(ph), z};
        double a[3] = {fabs(nrm[0]) < 0.9 ? 1.0 : 0.0, fabs(nrm[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        double t1[3] = {nrm[1] * a[2] - nrm[2] * a[1], nrm[2] * a[0] - nrm[0] * a[2],
                        nrm[0] * a[1] - nrm[1] * a[0]};
        double l1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
        for (int k = 0; k < 3; ++k) t1[k] /= l1;
        double t2[3] = {nrm[1] * t1[2] - nrm[2] * t1[1], nrm[2] * t1[0] - nrm[0] * t1[2],
                        nrm[0] * t1[1] - nrm[1] * t1[0]};
        double th0 = 6.283185307179586 * rng.uniform();
        double R = L / sqrt(3.0);  // circumradius of an equilateral triangle
        for (int v = 0; v < 3; ++v)
        {
            double th = th0 + v * 2.0943951023931953;
            for (int k = 0; k < 3; ++k)
                tris9[9 * (size_t)i + 3 * v + k] =
                    (float)(c[k] + R * (cos(th) * t1[k] + sin(th) * t2[k]));
        }
    }
    return SPG_OK;
}

// Texel (f, x, y) holds the sky at the texel-corner direction of
// CreateCubeMap (src/cubemap.cpp:259-275): norm(fwd + right*(2x/W-1) +
// up*(2(1-y/H)-1)) with the face bases of src/cubemap.cpp:55-103.
int spg_gen_sky_cubemap(uint32_t size, const float sunDir[3],
                        float sunAngularRadius, float sunRadiance, float *rgba6)
{
    Sky sky;
    if (!rgba6 || !sunDir || size == 0 ||
        !make_sky(sunDir, sunAngularRadius, sunRadiance, sky))
        return fail(SPG_E_INVALID, "spg_gen_sky_cubemap: bad argument");
    static const double B[6][9] = {
        {1, 0, 0, 0, 1, 0, 0, 0, -1}, {-1, 0, 0, 0, 1, 0, 0, 0, 1},
        {0, 1, 0, 0, 0, -1, 1, 0, 0}, {0, -1, 0, 0, 0, 1, 1, 0, 0},
        {0, 0, 1, 0, 1, 0, 1, 0, 0},  {0, 0, -1, 0, 1, 0, -1, 0, 0}};
    for (uint32_t f = 0; f < 6; ++f)
        for (uint32_t y = 0; y < size; ++y)
            for (uint32_t x = 0; x < size; ++x)
            {
                double fx = 2.0 * x / size - 1.0, fy = 2.0 * (1.0 - (double)y / size) - 1.0;
                double d[3];
                for (int k = 0; k < 3; ++k)
                    d[k] = B[f][k] + B[f][6 + k] * fx + B[f][3 + k] * fy;
                double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                for (int k = 0; k < 3; ++k) d[k] /= l;
                float *px = rgba6 + 4 * (((size_t)f * size + y) * size + x);
                sky.eval(d, px);
                px[3] = 1.0f;
            }
    return SPG_OK;
}

// Texel (x, y) holds the sky at the direction the reference's equirect
// lookup maps to its centre (src/math_lib.h:849-884, uv.y flipped as in
// src/sp_material_system.cpp:88-93): azimuth 2 pi u, cos(inclination) = 1-2v.
int spg_gen_sky_equirect(uint32_t width, uint32_t height, const float sunDir[3],
                         float sunAngularRadius, float sunRadiance, float *rgba)
{
    Sky sky;
    if (!rgba || !sunDir || width == 0 || height == 0 ||
        !make_sky(sunDir, sunAngularRadius, sunRadiance, sky))
        return fail(SPG_E_INVALID, "spg_gen_sky_equirect: bad argument");
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x)
        {
            double u = (x + 0.5) / width, v = (y + 0.5) / height;
            double az = 6.283185307179586 * u, ci = 1.0 - 2.0 * v;
            double si = sqrt(std::max(0.0, 1.0 - ci * ci));
            double d[3] = {si * cos(az), ci, si * sin(az)};
            float *px = rgba + 4 * ((size_t)y * width + x);
            sky.eval(d, px);
            px[3] = 1.0f;
        }
    return SPG_OK;
}

}  // extern "C"
