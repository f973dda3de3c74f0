// C-ABI implementation: device and scene lifecycle, SAH builds, HBM upload
// and the wavefront trace driver.  Kernels live in spg_trace.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "spg_internal.h"
#include "spg_env_table.h"

namespace spg {

int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return SPG_OK;
    return fail(SPG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- host mat4 arithmetic in the reference's order (src/math_lib.h) ----
struct M4 { float m[16]; };
static M4 m4zero() { M4 r; memset(r.m, 0, 64); return r; }
static M4 m4identity()
{
    M4 r = m4zero();
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}
static M4 m4scale(v3 s)
{
    M4 r = m4zero();
    r.m[0] = s.x; r.m[5] = s.y; r.m[10] = s.z; r.m[15] = 1.0f;
    return r;
}
static M4 m4translate(v3 t)
{
    M4 r = m4identity();
    r.m[12] = t.x; r.m[13] = t.y; r.m[14] = t.z;
    return r;
}
// operator*(mat4, mat4), src/math_lib.h:250-270
static M4 m4mul(const M4 &a, const M4 &b)
{
    M4 r = m4zero();
    for (int i = 0; i < 4; ++i)
    {
        float row[4] = {a.m[i], a.m[4 + i], a.m[8 + i], a.m[12 + i]};
        for (int c = 0; c < 4; ++c)
            r.m[4 * c + i] = row[0] * b.m[4 * c] + row[1] * b.m[4 * c + 1] +
                             row[2] * b.m[4 * c + 2] + row[3] * b.m[4 * c + 3];
    }
    return r;
}
struct Q { float x, y, z, w; };
// quat operator*, src/math_lib.h:565-571
static Q qmul(Q p, Q q)
{
    v3 pv{p.x, p.y, p.z}, qv{q.x, q.y, q.z};
    v3 v = (qv * p.w + pv * q.w) + cross(pv, qv);
    return Q{v.x, v.y, v.z, p.w * q.w - dot(pv, qv)};
}
static Q qconj(Q p) { return Q{-p.x, -p.y, -p.z, p.w}; }
// RotateVector, src/math_lib.h:588-593
static v3 qrotate(v3 v, Q p)
{
    Q q{v.x, v.y, v.z, 0.0f};
    Q r = qmul(qmul(p, q), qconj(p));
    return v3{r.x, r.y, r.z};
}
// Rotate(quat), src/math_lib.h:602-623
static M4 m4rotate(Q q)
{
    M4 r = m4identity();
    float x = q.x, y = q.y, z = q.z, w = q.w;
    r.m[0] = 1.0f - 2.0f * y * y - 2.0f * z * z;
    r.m[1] = 2.0f * x * y + 2.0f * z * w;
    r.m[2] = 2.0f * x * z - 2.0f * y * w;
    r.m[4] = 2.0f * x * y - 2.0f * z * w;
    r.m[5] = 1.0f - 2.0f * x * x - 2.0f * z * z;
    r.m[6] = 2.0f * y * z + 2.0f * x * w;
    r.m[8] = 2.0f * x * z + 2.0f * y * w;
    r.m[9] = 2.0f * y * z - 2.0f * x * w;
    r.m[10] = 1.0f - 2.0f * x * x - 2.0f * y * y;
    return r;
}
static v3 m4point(const M4 &a, v3 p)
{
    float r[3];
    for (int i = 0; i < 3; ++i)
        r[i] = a.m[i] * p.x + a.m[4 + i] * p.y + a.m[8 + i] * p.z +
               a.m[12 + i] * 1.0f;
    return v3{r[0], r[1], r[2]};
}
// Sign of the determinant of a column-major 4x4's upper 3x3, evaluated in
// double as col0 . (col1 x col2); 0 when it is 0 or not finite.
static int32_t model_orient(const float *m)
{
    const double x = (double)m[5] * m[10] - (double)m[6] * m[9];
    const double y = (double)m[6] * m[8] - (double)m[4] * m[10];
    const double z = (double)m[4] * m[9] - (double)m[5] * m[8];
    const double det = (double)m[0] * x + (double)m[1] * y + (double)m[2] * z;
    return det > 0.0 ? 1 : det < 0.0 ? -1 : 0;
}

static m34 rows34(const float *m)
{
    m34 r;
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 4; ++c) r.r[i][c] = m[4 * c + i];
    return r;
}

template <typename T>
static int upload(void **dst, const T *src, size_t count)
{
    if (*dst) hipFree(*dst);
    *dst = nullptr;
    size_t bytes = std::max<size_t>(sizeof(T) * count, 16);
    int r = hip_check(hipMalloc(dst, bytes), "hipMalloc");
    if (r) return r;
    if (count) return hip_check(hipMemcpy(*dst, src, sizeof(T) * count,
                                          hipMemcpyHostToDevice),
                                "hipMemcpy");
    return SPG_OK;
}

static void free_ptr(void *&p)
{
    if (p) hipFree(p);
    p = nullptr;
}

}  // namespace spg

using namespace spg;


extern "C" {

uint32_t spg_pixel_seed(uint32_t base, uint32_t index)
{
    return pixel_seed(base, index);
}

// ------------------------------------------------------------ lifecycle
int spg_device_create(int ordinal, spg_Device **out)
{
    if (!out) return fail(SPG_E_INVALID, "spg_device_create: out is NULL");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (ordinal < 0 || ordinal >= n)
        return fail(SPG_E_INVALID, "spg_device_create: no device " +
                                       std::to_string(ordinal));
    HIPCHK(hipSetDevice(ordinal));
    spg_Device *d = new spg_Device();
    d->ordinal = ordinal;
    hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
    {
        delete d;
        return hip_check(e, "hipStreamCreate");
    }
    for (int i = 0; i < kEvents; ++i) hipEventCreate(&d->ev[i]);
    d->eventsMade = true;
    // persistent-launch work counters (256 B) then the traversal stack spill
    size_t spillBytes = spill_ints() * sizeof(int);
    e = hipMalloc(&d->persistBlock, 256 + spillBytes);
    if (e != hipSuccess)
    {
        for (int i = 0; i < kEvents; ++i) hipEventDestroy(d->ev[i]);
        hipStreamDestroy(d->stream);
        delete d;
        return hip_check(e, "hipMalloc(traversal spill)");
    }
    d->persist.counter = (uint32_t *)d->persistBlock;
    d->persist.spill = (int *)((char *)d->persistBlock + 256);
    // the allocation-free entry point's scratch: this device's copy of the
    // static array, looked up once with the device current
    e = noalloc_scratch_address(&d->noallocScratch);
    if (e != hipSuccess)
    {
        for (int i = 0; i < kEvents; ++i) hipEventDestroy(d->ev[i]);
        hipStreamDestroy(d->stream);
        hipFree(d->persistBlock);
        delete d;
        return hip_check(e, "hipGetSymbolAddress(scratch)");
    }
    *out = d;
    return SPG_OK;
}

static void free_paths(spg_Device *d)
{
    free_ptr(d->pathsBlock);
    d->pathsBytes = 0;
    memset(&d->paths, 0, sizeof d->paths);
}

static void delete_device(spg_Device *d)
{
    hipSetDevice(d->ordinal);
    hipStreamSynchronize(d->stream);
    free_paths(d);
    free_ptr(d->image);
    if (d->hostStage) hipHostFree(d->hostStage);
    free_ptr(d->slotPixel);
    free_ptr(d->accum);
    free_ptr(d->persistBlock);
    if (d->eventsMade)
        for (int i = 0; i < kEvents; ++i) hipEventDestroy(d->ev[i]);
    hipStreamDestroy(d->stream);
    delete d;
}

// A device outlives its scenes: destroying it while scenes are alive defers
// the release to the last spg_scene_destroy.
int spg_device_destroy(spg_Device *d)
{
    if (!d) return SPG_OK;
    {
        std::lock_guard<std::mutex> g(d->lock);
        if (d->liveScenes > 0)
        {
            d->destroyRequested = true;
            return SPG_OK;
        }
    }
    delete_device(d);
    return SPG_OK;
}

int spg_device_synchronize(spg_Device *d)
{
    if (!d) return fail(SPG_E_INVALID, "device is NULL");
    HIPCHK(hipSetDevice(d->ordinal));
    return hip_check(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
}

int spg_device_alloc(spg_Device *d, size_t bytes, void **out)
{
    if (!d || !out) return fail(SPG_E_INVALID, "spg_device_alloc: NULL");
    HIPCHK(hipSetDevice(d->ordinal));
    return hip_check(hipMalloc(out, bytes ? bytes : 16), "hipMalloc");
}

int spg_device_free(spg_Device *d, void *p)
{
    if (!d) return fail(SPG_E_INVALID, "spg_device_free: NULL device");
    HIPCHK(hipSetDevice(d->ordinal));
    return hip_check(hipFree(p), "hipFree");
}

int spg_copy_to_device(spg_Device *d, void *dst, const void *src, size_t bytes)
{
    if (!d) return fail(SPG_E_INVALID, "NULL device");
    HIPCHK(hipSetDevice(d->ordinal));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, d->stream));
    return hip_check(hipStreamSynchronize(d->stream), "sync");
}

int spg_copy_to_host(spg_Device *d, void *dst, const void *src, size_t bytes)
{
    if (!d) return fail(SPG_E_INVALID, "NULL device");
    HIPCHK(hipSetDevice(d->ordinal));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, d->stream));
    return hip_check(hipStreamSynchronize(d->stream), "sync");
}

// ---------------------------------------------------------------- scene
int spg_scene_create(spg_Device *dev, spg_Scene **out)
{
    if (!dev || !out) return fail(SPG_E_INVALID, "spg_scene_create: NULL");
    std::lock_guard<std::mutex> g(dev->lock);
    if (dev->destroyRequested)
        return fail(SPG_E_STATE, "spg_scene_create: device is being destroyed");
    spg_Scene *s = new spg_Scene();
    s->dev = dev;
    dev->liveScenes++;
    *out = s;
    return SPG_OK;
}

static void free_scene_device(spg_Scene *s)
{
    for (auto &m : s->meshes)
    {
        free_ptr(m.dNodes);
        m.dTris = nullptr;
        free_ptr(m.dVerts);
        free_ptr(m.dIdx);
        free_ptr(m.dShade);
    }
    free_ptr(s->dTlasNodes);
    free_ptr(s->dInst);
    free_ptr(s->dMeshes);
    free_ptr(s->dMats);
    free_ptr(s->dTexs);
    free_ptr(s->dTexels);
    free_ptr(s->dEnv);
    free_ptr(s->dEmit);
}

int spg_scene_destroy(spg_Scene *s)
{
    if (!s) return SPG_OK;
    spg_Device *d = s->dev;
    bool last;
    {
        std::lock_guard<std::mutex> g(d->lock);
        hipSetDevice(d->ordinal);
        hipStreamSynchronize(d->stream);
        free_scene_device(s);
        delete s;
        d->liveScenes--;
        last = d->destroyRequested && d->liveScenes == 0;
    }
    if (last) delete_device(d);
    return SPG_OK;
}

int spg_mesh_create(spg_Scene *s, const float *verts8, uint32_t nv,
                    const uint32_t *idx, uint32_t ni, int smooth,
                    uint32_t *meshId)
{
    if (!s || (!verts8 && nv) || (!idx && ni))
        return fail(SPG_E_INVALID, "spg_mesh_create: NULL argument");
    if (ni % 3 != 0 || nv == 0 || ni == 0)
        return fail(SPG_E_INVALID,
                    "spg_mesh_create: need nv > 0 and a positive multiple of 3 indices");
    for (uint32_t i = 0; i < ni; ++i)
        if (idx[i] >= nv)
            return fail(SPG_E_INVALID, "spg_mesh_create: index out of range");
    // finite positions keep every BVH box finite, which the traversal's
    // min/max slab test relies on
    for (uint32_t i = 0; i < nv; ++i)
        if (!std::isfinite(verts8[8 * i]) || !std::isfinite(verts8[8 * i + 1]) ||
            !std::isfinite(verts8[8 * i + 2]))
            return fail(SPG_E_INVALID, "spg_mesh_create: non-finite vertex position");
    std::lock_guard<std::mutex> g(s->dev->lock);
    s->meshes.emplace_back();
    HostMesh &m = s->meshes.back();
    m.verts.assign(verts8, verts8 + (size_t)nv * 8);
    m.idx.assign(idx, idx + ni);
    m.smooth = smooth ? 1 : 0;
    // ComputeAabb over all vertices (src/sp_scene.cpp:56-74)
    v3 lo{verts8[0], verts8[1], verts8[2]}, hi = lo;
    for (uint32_t i = 1; i < nv; ++i)
    {
        v3 p{verts8[8 * i], verts8[8 * i + 1], verts8[8 * i + 2]};
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    memcpy(m.boundsLo, &lo, 12);
    memcpy(m.boundsHi, &hi, 12);
    s->built = false;
    if (meshId) *meshId = (uint32_t)(s->meshes.size() - 1);
    return SPG_OK;
}

int spg_scene_add_object(spg_Scene *s, uint32_t meshId, uint32_t materialId,
                         const float pos[3], const float quat[4],
                         const float scale[3], uint32_t *objectId)
{
    if (!s || !pos || !quat || !scale)
        return fail(SPG_E_INVALID, "spg_scene_add_object: NULL argument");
    for (int k = 0; k < 4; ++k)
        if ((k < 3 && (!std::isfinite(pos[k]) || !std::isfinite(scale[k]))) ||
            !std::isfinite(quat[k]))
            return fail(SPG_E_INVALID, "spg_scene_add_object: non-finite transform");
    std::lock_guard<std::mutex> g(s->dev->lock);
    if (meshId >= s->meshes.size())
        return fail(SPG_E_NOT_FOUND, "spg_scene_add_object: unknown mesh");
    HostObject o;
    o.mesh = meshId;
    o.materialId = materialId;
    memcpy(o.pos, pos, 12);
    memcpy(o.quat, quat, 16);
    memcpy(o.scale, scale, 12);
    v3 P{pos[0], pos[1], pos[2]}, S{scale[0], scale[1], scale[2]};
    Q qq{quat[0], quat[1], quat[2], quat[3]};
    // sp_AddObjectToScene, src/sp_scene.cpp:77-117
    M4 model = m4mul(m4mul(m4translate(P), m4rotate(qq)), m4scale(S));
    M4 inv = m4mul(m4mul(m4scale(inverse(S)), m4rotate(qconj(qq))),
                   m4translate(-P));
    memcpy(o.model, model.m, 64);
    memcpy(o.inv, inv.m, 64);
    // TransformAabb, src/aabb.h:29-57
    const HostMesh &m = s->meshes[meshId];
    v3 lo, hi;
    memcpy(&lo, m.boundsLo, 12);
    memcpy(&hi, m.boundsHi, 12);
    v3 c[8] = {{lo.x, lo.y, lo.z}, {hi.x, lo.y, lo.z}, {hi.x, lo.y, hi.z},
               {lo.x, lo.y, hi.z}, {lo.x, hi.y, lo.z}, {hi.x, hi.y, lo.z},
               {hi.x, hi.y, hi.z}, {lo.x, hi.y, hi.z}};
    v3 rmin = m4point(model, c[0]), rmax = rmin;
    for (int i = 1; i < 8; ++i)
    {
        v3 p = m4point(model, c[i]);
        rmin = vmin(rmin, p);
        rmax = vmax(rmax, p);
    }
    memcpy(o.aabbLo, &rmin, 12);
    memcpy(o.aabbHi, &rmax, 12);
    // The instance tree bounds each instance by its transformed vertices
    // (the box of a rotated mesh box can be up to sqrt(3) wider per axis):
    // fewer rays enter instances they miss.  The result does not depend on
    // it, since the tree's padded boxes contain every triangle either way.
    v3 tmin = rmin, tmax = rmax;
    if (!m.verts.empty())
    {
        tmin = tmax = m4point(model, v3{m.verts[0], m.verts[1], m.verts[2]});
        for (size_t i = 8; i < m.verts.size(); i += 8)
        {
            v3 p = m4point(model, v3{m.verts[i], m.verts[i + 1], m.verts[i + 2]});
            tmin = vmin(tmin, p);
            tmax = vmax(tmax, p);
        }
    }
    memcpy(o.tightLo, &tmin, 12);
    memcpy(o.tightHi, &tmax, 12);
    s->objects.push_back(o);
    s->built = false;
    if (objectId) *objectId = (uint32_t)(s->objects.size() - 1);
    return SPG_OK;
}

static int build_mesh(HostMesh &m, const spg_BuildOpts &opt)
{
    uint32_t tc = (uint32_t)(m.idx.size() / 3);
    std::vector<float> lo(3 * (size_t)tc), hi(3 * (size_t)tc);
    for (uint32_t i = 0; i < tc; ++i)
    {
        const float *a = &m.verts[8 * (size_t)m.idx[3 * i]];
        const float *b = &m.verts[8 * (size_t)m.idx[3 * i + 1]];
        const float *c = &m.verts[8 * (size_t)m.idx[3 * i + 2]];
        for (int k = 0; k < 3; ++k)
        {
            lo[3 * i + k] = fmin_sel(a[k], fmin_sel(b[k], c[k]));
            hi[3 * i + k] = fmax_sel(a[k], fmax_sel(b[k], c[k]));
        }
    }
    BvhBuildInput in{lo.data(), hi.data(), tc, opt.maxLeafSize, opt.sahBins,
                     opt.threads, kMaxBlasDepth};
    if (!build_bvh(in, m.bvh))
        return fail(SPG_E_INVALID, "mesh BVH build failed");
    if (m.bvh.maxDepth > kMaxBlasDepth)
        return fail(SPG_E_CAPACITY, "mesh BVH deeper than the traversal stack");
    if (!collapse_wide(m.bvh, m.wide, m.widePending))
        return fail(SPG_E_INVALID, "mesh BVH quantization failed");
    if (m.widePending > 3 * (kMaxBlasDepth + 1))
        return fail(SPG_E_CAPACITY, "mesh hierarchy deeper than the traversal stack");
    // One allocation of 64-byte records: the wide nodes, then the triangles
    // in leaf order as (a.xyz, e1.x) (e1.yz, e2.xy) (e2.z, prim, n.xy) (n.z,
    // 0, 0, 0) with e1 = b - a, e2 = c - a, n = cross(e1, e2) (the
    // Moller-Trumbore quantities, src/ray_intersection.cpp:156-190, formed
    // with the same FP32 operations as the test).  Triangle k is record
    // nn + k, so the device copy's leaf references name records directly
    // and a traversal step fetches a node or a triangle alike.
    const size_t nn = m.wide.size();
    if (nn + (size_t)tc >= (size_t)1 << 28)
        return fail(SPG_E_CAPACITY, "mesh too large for 28-bit leaf references");
    std::vector<float4> blob(4 * (nn + (size_t)tc), make_float4(0, 0, 0, 0));
    std::vector<WideNode> nodes = m.wide;
    for (WideNode &w : nodes)
        for (int32_t &r : w.ref)
            if (r != kEmptyChild && ref_is_leaf(r))
                r = leaf_ref((uint32_t)nn + leaf_first(r), leaf_count(r));
    memcpy(blob.data(), nodes.data(), nn * sizeof(WideNode));
    float4 *tris = blob.data() + 4 * nn;
    for (uint32_t k = 0; k < tc; ++k)
    {
        uint32_t prim = m.bvh.prims[k];
        const float *a = &m.verts[8 * (size_t)m.idx[3 * prim]];
        const float *b = &m.verts[8 * (size_t)m.idx[3 * prim + 1]];
        const float *c = &m.verts[8 * (size_t)m.idx[3 * prim + 2]];
        v3 e1{b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        v3 e2{c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        v3 n = cross(e1, e2);
        float primBits;
        memcpy(&primBits, &prim, 4);
        tris[4 * k] = make_float4(a[0], a[1], a[2], e1.x);
        tris[4 * k + 1] = make_float4(e1.y, e1.z, e2.x, e2.y);
        tris[4 * k + 2] = make_float4(e2.z, primBits, n.x, n.y);
        tris[4 * k + 3] = make_float4(n.z, 0.0f, 0.0f, 0.0f);
    }
    CHK(upload(&m.dNodes, blob.data(), blob.size()));
    m.dTris = (float4 *)m.dNodes + 4 * nn;
    m.nodeCount = (uint32_t)nn;
    CHK(upload(&m.dVerts, m.verts.data(), m.verts.size()));
    CHK(upload(&m.dIdx, m.idx.data(), m.idx.size()));
    // shading records by triangle index, one 64-byte fetch at a hit instead
    // of the index and three vertex reads: (n0.xyz, n1.x) (n1.yz, n2.xy)
    // (n2.z, uv0, uv1.x) (uv1.y, uv2, 0); a flat mesh stores
    // cross(b - a, c - a) (the same FP32 operations as the kernel's) as n0
    std::vector<float4> shade(4 * (size_t)tc);
    for (uint32_t k = 0; k < tc; ++k)
    {
        const float *p0 = &m.verts[8 * (size_t)m.idx[3 * k]];
        const float *p1 = &m.verts[8 * (size_t)m.idx[3 * k + 1]];
        const float *p2 = &m.verts[8 * (size_t)m.idx[3 * k + 2]];
        v3 n0{p0[3], p0[4], p0[5]}, n1{p1[3], p1[4], p1[5]}, n2{p2[3], p2[4], p2[5]};
        if (!m.smooth)
        {
            v3 a{p0[0], p0[1], p0[2]}, b{p1[0], p1[1], p1[2]}, c{p2[0], p2[1], p2[2]};
            n0 = cross(b - a, c - a);
            n1 = n2 = v3{0.0f, 0.0f, 0.0f};
        }
        shade[4 * (size_t)k] = make_float4(n0.x, n0.y, n0.z, n1.x);
        shade[4 * (size_t)k + 1] = make_float4(n1.y, n1.z, n2.x, n2.y);
        shade[4 * (size_t)k + 2] = make_float4(n2.z, p0[6], p0[7], p1[6]);
        shade[4 * (size_t)k + 3] = make_float4(p1[7], p2[6], p2[7], 0.0f);
    }
    CHK(upload(&m.dShade, shade.data(), shade.size()));
    m.built = true;
    return SPG_OK;
}

static int find_index(const std::vector<uint32_t> &keys, uint32_t id)
{
    for (size_t i = 0; i < keys.size(); ++i)
        if (keys[i] == id) return (int)i;
    return -1;
}

// Uploads the material and texture tables and the instance records (which
// hold resolved material slots), as sp_FindMaterialById/sp_FindTexture would
// resolve them (first registration of an id wins).
static int upload_tables(spg_Scene *s)
{
    std::vector<uint32_t> texKeys;
    for (auto &t : s->textures) texKeys.push_back(t.id);
    std::vector<DevTexture> texs;
    std::vector<float> texels;
    for (auto &t : s->textures)
    {
        DevTexture d{};
        d.offset = texels.size() / 4;
        d.width = t.width;
        d.height = t.height;
        d.isCube = t.isCube;
        texels.insert(texels.end(), t.rgba.begin(), t.rgba.end());
        texs.push_back(d);
    }
    std::vector<DevMaterial> mats;
    for (auto &m : s->materials)
    {
        DevMaterial d{};
        memcpy(d.albedo, m.albedo, 12);
        memcpy(d.emission, m.emission, 12);
        d.roughness = m.roughness;
        d.albedoTex = find_index(texKeys, m.albedoTexture);
        d.emissionTex = find_index(texKeys, m.emissionTexture);
        mats.push_back(d);
    }
    std::vector<DevInstance> inst;
    for (auto &o : s->objects)
    {
        DevInstance d{};
        d.model = rows34(o.model);
        d.inv = rows34(o.inv);
        float smin = std::min(fabsf(o.scale[0]),
                              std::min(fabsf(o.scale[1]), fabsf(o.scale[2])));
        d.cullScale = smin > 0.0f ? 1.0f / smin : INFINITY;
        d.mesh = o.mesh;
        d.material = find_index(s->materialKeys, o.materialId);
        d.materialId = o.materialId;
        d.nodes = (const float4 *)s->meshes[o.mesh].dNodes;
        d.nodeCount = s->meshes[o.mesh].nodeCount;
        d.orient = model_orient(o.model);
        inst.push_back(d);
    }
    CHK(upload(&s->dTexs, texs.data(), texs.size()));
    CHK(upload(&s->dTexels, reinterpret_cast<const float4 *>(texels.data()),
               texels.size() / 4));
    CHK(upload(&s->dMats, mats.data(), mats.size()));
    CHK(upload(&s->dInst, inst.data(), inst.size()));
    DevScene &ds = s->dev_scene;
    ds.texs = (const DevTexture *)s->dTexs;
    ds.texels = (const float4 *)s->dTexels;
    ds.mats = (const DevMaterial *)s->dMats;
    ds.inst = (const DevInstance *)s->dInst;
    ds.background = find_index(s->materialKeys, s->backgroundId);
    // next-event estimation table of the background's cube-map emission
    ds.envPdf = ds.envProb = nullptr;
    ds.envAlias = nullptr;
    ds.envTex = -1;
    ds.envCount = 0;
    free_ptr(s->dEnv);
    if (ds.background >= 0 && mats[ds.background].emissionTex >= 0)
    {
        const int32_t ti = mats[ds.background].emissionTex;
        const HostTexture &t = s->textures[ti];
        std::vector<float> pdf, prob;
        std::vector<uint32_t> alias;
        if (t.isCube && build_env_table(t.rgba.data(), t.width, t.height, pdf, prob, alias))
        {
            const size_t n = pdf.size();
            HIPCHK(hipMalloc(&s->dEnv, 12 * n));
            char *e = (char *)s->dEnv;
            HIPCHK(hipMemcpy(e, pdf.data(), 4 * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + 4 * n, prob.data(), 4 * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + 8 * n, alias.data(), 4 * n, hipMemcpyHostToDevice));
            ds.envPdf = (const float *)e;
            ds.envProb = (const float *)(e + 4 * n);
            ds.envAlias = (const uint32_t *)(e + 8 * n);
            ds.envTex = ti;
            ds.envCount = (uint32_t)n;
        }
    }
    // next-event estimation table of the emissive triangles
    ds.emitPdfA = ds.emitProb = nullptr;
    ds.emitAlias = ds.emitPair = nullptr;
    ds.emitBase = nullptr;
    ds.emitCount = 0;
    ds.emitInst = nullptr;
    ds.emitSlot = nullptr;
    ds.emitProbL = nullptr;
    ds.emitAliasL = nullptr;
    ds.emitInstCount = 0;
    free_ptr(s->dEmit);
    {
        std::vector<EmitterInput> ein;
        for (auto &o : s->objects)
        {
            EmitterInput e{o.model, nullptr, nullptr, 0, 0.0};
            const int slot = find_index(s->materialKeys, o.materialId);
            const HostMesh &m = s->meshes[o.mesh];
            if (slot >= 0 && mats[slot].emissionTex < 0)
            {
                e.verts = m.verts.data();
                e.idx = m.idx.data();
                e.triCount = (uint32_t)(m.idx.size() / 3);
                e.lum = ((double)mats[slot].emission[0] + (double)mats[slot].emission[1] +
                         (double)mats[slot].emission[2]) / 3.0;
            }
            ein.push_back(e);
        }
        std::vector<float> pdfA, prob;
        std::vector<uint32_t> alias, pair;
        std::vector<int32_t> base;
        if (build_emitter_table(ein, pdfA, prob, alias, pair, base))
        {
            std::vector<float> rec, probL;
            std::vector<int32_t> slot;
            std::vector<uint32_t> aliasL;
            build_emitter_instances(ein, base, rec, slot, probL, aliasL);
            const size_t n = pdfA.size(), ni = base.size(), ne = rec.size() / 8;
            // 16-byte records first, then the 4-byte arrays
            const size_t oRec = 0, oPdf = 32 * ne, oProb = oPdf + 4 * n,
                         oAlias = oProb + 4 * n, oPair = oAlias + 4 * n,
                         oBase = oPair + 8 * n, oSlot = oBase + 4 * ni,
                         oProbL = oSlot + 4 * ni, oAliasL = oProbL + 4 * n,
                         bytes = oAliasL + 4 * n;
            HIPCHK(hipMalloc(&s->dEmit, bytes));
            char *e = (char *)s->dEmit;
            if (ne) HIPCHK(hipMemcpy(e + oRec, rec.data(), 32 * ne, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + oPdf, pdfA.data(), 4 * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + oProb, prob.data(), 4 * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + oAlias, alias.data(), 4 * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + oPair, pair.data(), 8 * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + oBase, base.data(), 4 * ni, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + oSlot, slot.data(), 4 * ni, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + oProbL, probL.data(), 4 * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e + oAliasL, aliasL.data(), 4 * n, hipMemcpyHostToDevice));
            ds.emitInst = (const float4 *)(e + oRec);
            ds.emitPdfA = (const float *)(e + oPdf);
            ds.emitProb = (const float *)(e + oProb);
            ds.emitAlias = (const uint32_t *)(e + oAlias);
            ds.emitPair = (const uint32_t *)(e + oPair);
            ds.emitBase = (const int32_t *)(e + oBase);
            ds.emitSlot = (const int32_t *)(e + oSlot);
            ds.emitProbL = (const float *)(e + oProbL);
            ds.emitAliasL = (const uint32_t *)(e + oAliasL);
            ds.emitCount = (uint32_t)n;
            ds.emitInstCount = (uint32_t)ne;
        }
    }
    ds.envSelect = ds.envTex >= 0 ? (ds.emitCount ? 0.5f : 1.0f) : 0.0f;
    s->dirty = false;
    return SPG_OK;
}

int spg_scene_learn_light_split(spg_Scene *s, uint32_t enable)
{
    if (!s) return fail(SPG_E_INVALID, "spg_scene_learn_light_split: NULL");
    std::lock_guard<std::mutex> g(s->dev->lock);
    s->learnLight = enable != 0;
    if (!enable) s->lightStat[0] = s->lightStat[1] = 0;
    return SPG_OK;
}

int spg_scene_light_split(spg_Scene *s, float *skyFraction)
{
    if (!s || !skyFraction) return fail(SPG_E_INVALID, "spg_scene_light_split: NULL");
    std::lock_guard<std::mutex> g(s->dev->lock);
    *skyFraction = light_split(s->lightStat, -1.0f);
    return SPG_OK;
}

int spg_scene_build(spg_Scene *s, const spg_BuildOpts *opts)
{
    if (!s) return fail(SPG_E_INVALID, "spg_scene_build: NULL scene");
    spg_BuildOpts opt{4, 16, 0};
    if (opts) opt = *opts;
    if (opt.maxLeafSize == 0) opt.maxLeafSize = 4;
    if (opt.maxLeafSize > 8)
        return fail(SPG_E_INVALID, "spg_scene_build: maxLeafSize > 8");
    if (opt.sahBins == 0) opt.sahBins = 16;
    std::lock_guard<std::mutex> g(s->dev->lock);
    HIPCHK(hipSetDevice(s->dev->ordinal));
    for (auto &m : s->meshes)
        if (!m.built) CHK(build_mesh(m, opt));
    std::vector<DevMesh> dm;
    for (auto &m : s->meshes)
    {
        DevMesh d{};
        d.nodes = (const float4 *)m.dNodes;
        d.tris = (const float4 *)m.dTris;
        d.verts = (const float *)m.dVerts;
        d.idx = (const uint32_t *)m.dIdx;
        d.shade = (const float4 *)m.dShade;
        d.smooth = m.smooth;
        d.triCount = (uint32_t)(m.idx.size() / 3);
        dm.push_back(d);
    }
    CHK(upload(&s->dMeshes, dm.data(), dm.size()));
    // broadphase: sp_BuildSceneBroadphase over object boxes
    // (src/sp_scene.cpp:119-125), here a SAH BVH2 with one object per leaf
    s->tlas = BvhBuildResult();
    uint32_t no = (uint32_t)s->objects.size();
    if (no > 0)
    {
        std::vector<float> lo(3 * (size_t)no), hi(3 * (size_t)no);
        for (uint32_t i = 0; i < no; ++i)
        {
            memcpy(&lo[3 * i], s->objects[i].tightLo, 12);
            memcpy(&hi[3 * i], s->objects[i].tightHi, 12);
        }
        BvhBuildInput in{lo.data(), hi.data(), no, 1, opt.sahBins, opt.threads,
                         kMaxTlasDepth};
        build_bvh(in, s->tlas);
        if (s->tlas.maxDepth > kMaxTlasDepth)
            return fail(SPG_E_CAPACITY, "instance BVH too deep");
        std::vector<WideNode> nodes;
        if (!collapse_wide(s->tlas, nodes, s->tlasPending))
            return fail(SPG_E_INVALID, "instance BVH quantization failed");
        uint32_t meshPending = 0;
        for (auto &m : s->meshes) meshPending = std::max(meshPending, m.widePending);
        if (s->tlasPending + meshPending > (uint32_t)kStackDepth)
            return fail(SPG_E_CAPACITY, "scene hierarchy deeper than the traversal stack");
        // one instance per leaf: the device copy's leaf references name the
        // instance itself instead of its position in the leaf order
        for (WideNode &w : nodes)
            for (int32_t &r : w.ref)
                if (r != kEmptyChild && ref_is_leaf(r))
                    r = leaf_ref(s->tlas.prims[leaf_first(r)], 1);
        CHK(upload(&s->dTlasNodes, nodes.data(), nodes.size()));
    }
    DevScene &ds = s->dev_scene;
    ds.tlasNodes = (const float4 *)s->dTlasNodes;
    ds.meshes = (const DevMesh *)s->dMeshes;
    ds.instCount = no;
    CHK(upload_tables(s));
    s->built = true;
    return SPG_OK;
}

// The mesh build of a host triangle list (tris9 = a,b,c per triangle).
static int host_build(const char *who, const float *tris9, uint32_t n,
                      const spg_BuildOpts *opts, BvhBuildResult &b)
{
    spg_BuildOpts opt{4, 16, 0};
    if (opts) opt = *opts;
    if (opt.maxLeafSize == 0) opt.maxLeafSize = 4;
    if (opt.maxLeafSize > 8)
        return fail(SPG_E_INVALID, std::string(who) + ": maxLeafSize > 8");
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
        {
            const float *t = tris9 + 9 * (size_t)i;
            if (!std::isfinite(t[k]) || !std::isfinite(t[3 + k]) ||
                !std::isfinite(t[6 + k]))
                return fail(SPG_E_INVALID, std::string(who) + ": non-finite vertex");
            lo[3 * i + k] = fmin_sel(t[k], fmin_sel(t[3 + k], t[6 + k]));
            hi[3 * i + k] = fmax_sel(t[k], fmax_sel(t[3 + k], t[6 + k]));
        }
    BvhBuildInput in{lo.data(), hi.data(), n, opt.maxLeafSize,
                     opt.sahBins ? opt.sahBins : 16, opt.threads, kMaxBlasDepth};
    build_bvh(in, b);
    return SPG_OK;
}

int spg_bvh_build_host(const float *tris9, uint32_t n, const spg_BuildOpts *opts,
                       float *nodes16, uint32_t nodeCap, uint32_t *order,
                       uint32_t *nodeCount, uint32_t *maxDepth)
{
    if (!tris9 || n == 0 || !nodes16 || !order || !nodeCount)
        return fail(SPG_E_INVALID, "spg_bvh_build_host: bad argument");
    BvhBuildResult b;
    CHK(host_build("spg_bvh_build_host", tris9, n, opts, b));
    if (b.nodes.size() > nodeCap)
        return fail(SPG_E_CAPACITY, "spg_bvh_build_host: nodeCap too small");
    memcpy(nodes16, b.nodes.data(), 64 * b.nodes.size());
    memcpy(order, b.prims.data(), 4 * b.prims.size());
    *nodeCount = (uint32_t)b.nodes.size();
    if (maxDepth) *maxDepth = b.maxDepth;
    return SPG_OK;
}

int spg_bvh_build_wide_host(const float *tris9, uint32_t n, const spg_BuildOpts *opts,
                            uint8_t *nodes64, uint32_t nodeCap, uint32_t *order,
                            uint32_t *nodeCount, uint32_t *maxPending)
{
    if (!tris9 || n == 0 || !nodes64 || !order || !nodeCount)
        return fail(SPG_E_INVALID, "spg_bvh_build_wide_host: bad argument");
    BvhBuildResult b;
    CHK(host_build("spg_bvh_build_wide_host", tris9, n, opts, b));
    std::vector<WideNode> w;
    uint32_t pend = 0;
    if (!collapse_wide(b, w, pend))
        return fail(SPG_E_INVALID, "spg_bvh_build_wide_host: quantization failed");
    if (w.size() > nodeCap)
        return fail(SPG_E_CAPACITY, "spg_bvh_build_wide_host: nodeCap too small");
    memcpy(nodes64, w.data(), 64 * w.size());
    memcpy(order, b.prims.data(), 4 * b.prims.size());
    *nodeCount = (uint32_t)w.size();
    if (maxPending) *maxPending = pend;
    return SPG_OK;
}

int spg_mesh_bvh_info(spg_Scene *s, uint32_t meshId, uint32_t *nodes,
                      uint32_t *leaves, uint32_t *maxDepth, double *sahCost)
{
    if (!s || meshId >= s->meshes.size() || !s->meshes[meshId].built)
        return fail(SPG_E_STATE, "spg_mesh_bvh_info: mesh not built");
    const BvhBuildResult &b = s->meshes[meshId].bvh;
    if (nodes) *nodes = (uint32_t)b.nodes.size();
    if (leaves) *leaves = b.leafCount;
    if (maxDepth) *maxDepth = b.maxDepth;
    if (sahCost) *sahCost = b.sahCost;
    return SPG_OK;
}

int spg_mesh_bvh_export(spg_Scene *s, uint32_t meshId, float *nodes16,
                        uint32_t nodeCap, uint32_t *primOrder,
                        uint32_t primCap, int32_t *root)
{
    if (!s || meshId >= s->meshes.size() || !s->meshes[meshId].built)
        return fail(SPG_E_STATE, "spg_mesh_bvh_export: mesh not built");
    const BvhBuildResult &b = s->meshes[meshId].bvh;
    if (nodeCap < b.nodes.size() || primCap < b.prims.size())
        return fail(SPG_E_CAPACITY, "spg_mesh_bvh_export: buffers too small");
    memcpy(nodes16, b.nodes.data(), 64 * b.nodes.size());
    memcpy(primOrder, b.prims.data(), 4 * b.prims.size());
    if (root) *root = 0;
    return SPG_OK;
}

int spg_mesh_wide_export(spg_Scene *s, uint32_t meshId, uint8_t *nodes64,
                         uint32_t nodeCap, uint32_t *primOrder, uint32_t primCap,
                         uint32_t *nodeCount, uint32_t *primCount,
                         uint32_t *maxPending)
{
    if (!s || meshId >= s->meshes.size() || !s->meshes[meshId].built)
        return fail(SPG_E_STATE, "spg_mesh_wide_export: mesh not built");
    const HostMesh &m = s->meshes[meshId];
    if (nodeCount) *nodeCount = (uint32_t)m.wide.size();
    if (primCount) *primCount = (uint32_t)m.bvh.prims.size();
    if (maxPending) *maxPending = m.widePending;
    if (!nodes64 && !primOrder) return SPG_OK;  // size query
    if (!nodes64 || !primOrder || nodeCap < m.wide.size() ||
        primCap < m.bvh.prims.size())
        return fail(SPG_E_CAPACITY, "spg_mesh_wide_export: buffers too small");
    memcpy(nodes64, m.wide.data(), 64 * m.wide.size());
    memcpy(primOrder, m.bvh.prims.data(), 4 * m.bvh.prims.size());
    return SPG_OK;
}

// ------------------------------------------------------------ materials
int spg_material_register(spg_Scene *s, const spg_Material *m, uint32_t id)
{
    if (!s || !m) return fail(SPG_E_INVALID, "spg_material_register: NULL");
    std::lock_guard<std::mutex> g(s->dev->lock);
    if (s->materials.size() >= 4096)
        return fail(SPG_E_CAPACITY, "spg_material_register: table full");
    s->materialKeys.push_back(id);
    s->materials.push_back(*m);
    s->dirty = true;
    return SPG_OK;
}

static int add_texture(spg_Scene *s, const float *rgba, uint32_t w, uint32_t h,
                       uint32_t id, uint32_t cube)
{
    if (!s || !rgba || w == 0 || h == 0)
        return fail(SPG_E_INVALID, "texture: NULL pixels or zero size");
    std::lock_guard<std::mutex> g(s->dev->lock);
    if (s->textures.size() >= 256)
        return fail(SPG_E_CAPACITY, "texture table full");
    HostTexture t;
    t.id = id;
    t.width = w;
    t.height = h;
    t.isCube = cube;
    size_t n = (size_t)w * h * 4 * (cube ? 6 : 1);
    t.rgba.assign(rgba, rgba + n);
    s->textures.push_back(std::move(t));
    s->dirty = true;
    return SPG_OK;
}

int spg_texture_register(spg_Scene *s, const float *rgba, uint32_t w,
                         uint32_t h, uint32_t id)
{
    return add_texture(s, rgba, w, h, id, 0);
}

int spg_cubemap_register(spg_Scene *s, const float *rgba6, uint32_t w,
                         uint32_t h, uint32_t id)
{
    return add_texture(s, rgba6, w, h, id, 1);
}

int spg_set_background_material(spg_Scene *s, uint32_t id)
{
    if (!s) return fail(SPG_E_INVALID, "NULL scene");
    std::lock_guard<std::mutex> g(s->dev->lock);
    s->backgroundId = id;
    s->dirty = true;
    return SPG_OK;
}

// ---------------------------------------------------------------- camera
int spg_camera_configure(spg_Camera *c, uint32_t w, uint32_t h,
                         const float pos[3], const float quat[4],
                         float filmDistance)
{
    if (!c || !pos || !quat || w == 0 || h == 0)
        return fail(SPG_E_INVALID, "spg_camera_configure: bad argument");
    // sp_ConfigureCamera, src/simd_path_tracer.cpp:1-36
    Q q{quat[0], quat[1], quat[2], quat[3]};
    v3 P{pos[0], pos[1], pos[2]};
    v3 right = qrotate(v3{1, 0, 0}, q);
    v3 up = qrotate(v3{0, 1, 0}, q);
    v3 fwd = qrotate(v3{0, 0, -1}, q);
    v3 fc = P + fwd * filmDistance;
    memcpy(c->right, &right, 12);
    memcpy(c->up, &up, 12);
    memcpy(c->forward, &fwd, 12);
    memcpy(c->position, &P, 12);
    memcpy(c->filmCenter, &fc, 12);
    c->width = w;
    c->height = h;
    c->halfPixelWidth = 0.5f / (float)w;
    c->halfPixelHeight = 0.5f / (float)h;
    float filmWidth = 1.0f, filmHeight = 1.0f;
    if (w > h) filmHeight = (float)h / (float)w;
    else if (w < h) filmWidth = (float)w / (float)h;
    c->halfFilmWidth = 0.5f * filmWidth;
    c->halfFilmHeight = 0.5f * filmHeight;
    return SPG_OK;
}

// ComputeTiles, src/tile.h:11-42
int spg_compute_tiles(uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
                      spg_Tile *tiles, uint32_t maxTiles, uint32_t *count)
{
    if (tw == 0 || th == 0 || (!tiles && maxTiles) || !count)
        return fail(SPG_E_INVALID, "spg_compute_tiles: bad argument");
    uint32_t ny = (uint32_t)ceilf((float)H / (float)th);
    uint32_t nx = (uint32_t)ceilf((float)W / (float)tw);
    for (uint32_t ty = 0; ty < ny; ++ty)
        for (uint32_t tx = 0; tx < nx; ++tx)
        {
            uint32_t index = tx + ty * nx;
            if (index >= maxTiles) break;
            spg_Tile t;
            t.minX = tx * tw;
            t.minY = ty * th;
            t.maxX = std::min(t.minX + tw, W);
            t.maxY = std::min(t.minY + th, H);
            tiles[index] = t;
        }
    *count = std::min(ny * nx, maxTiles);
    return SPG_OK;
}

}  // extern "C"

// ------------------------------------------------------------ tracing
namespace spg {

DevCamera dev_camera(const spg_Camera &c)
{
    DevCamera r{};
    memcpy(&r.right, c.right, 12);
    memcpy(&r.up, c.up, 12);
    memcpy(&r.forward, c.forward, 12);
    memcpy(&r.position, c.position, 12);
    memcpy(&r.filmCenter, c.filmCenter, 12);
    r.width = c.width;
    r.height = c.height;
    r.halfPixelWidth = c.halfPixelWidth;
    r.halfPixelHeight = c.halfPixelHeight;
    r.halfFilmWidth = c.halfFilmWidth;
    r.halfFilmHeight = c.halfFilmHeight;
    return r;
}

int scene_sync(spg_Scene *s)
{
    if (!s->built) return fail(SPG_E_STATE, "scene not built");
    if (s->dirty) CHK(upload_tables(s));
    return SPG_OK;
}

// Wavefront state for `capacity` paths of up to `bounces` rays, one
// allocation (hipFree of `block` releases it).
int alloc_path_state(DevPaths &p, void *&block, size_t &bytes, uint32_t capacity,
                     uint32_t bounces)
{
    capacity = std::max(capacity, 1024u);
    size_t P = capacity;
    size_t off = 0;
    auto take = [&](size_t b) {
        size_t o = off;
        off += (b + 255) / 256 * 256;
        return o;
    };
    size_t oSlot = take(4 * P), oSample = take(4 * P), oXs = take(4 * P),
           oPcg = take(8 * P), oSlotRng = take(4 * P),
           oVtx = take(4 * P * 7 * bounces), oVc = take(4 * P),
           oBeta = take(12 * P), oRad = take(12 * P), oQo0 = take(12 * P),
           oQo1 = take(12 * P), oQd0 = take(12 * P), oQd1 = take(12 * P),
           oQp0 = take(4 * P), oQp1 = take(4 * P), oQo2 = take(12 * P),
           oQd2 = take(12 * P), oQp2 = take(4 * P), oPdf = take(4 * P),
           oLastN = take(12 * P), oLastW = take(4 * P),
           oNee = take(12 * P), oTgt = take(8 * P), oDist = take(4 * P),
           oCnt = take(64),
           oMet = take(128), oHt = take(4 * P), oHu = take(4 * P),
           oHv = take(4 * P), oHi = take(4 * P), oHp = take(4 * P);
    HIPCHK(hipMalloc(&block, off));
    char *b = (char *)block;
    p.capacity = (uint32_t)P;
    p.bounces = bounces;
    p.pathSlot = (uint32_t *)(b + oSlot);
    p.pathSample = (uint32_t *)(b + oSample);
    p.rngXs = (uint32_t *)(b + oXs);
    p.rngPcg = (uint64_t *)(b + oPcg);
    p.slotRng = (uint32_t *)(b + oSlotRng);
    p.vtx = (float *)(b + oVtx);
    p.vcount = (uint32_t *)(b + oVc);
    p.beta = (float *)(b + oBeta);
    p.radiance = (float *)(b + oRad);
    p.qo[0] = (float *)(b + oQo0);
    p.qo[1] = (float *)(b + oQo1);
    p.qd[0] = (float *)(b + oQd0);
    p.qd[1] = (float *)(b + oQd1);
    p.qpath[0] = (uint32_t *)(b + oQp0);
    p.qpath[1] = (uint32_t *)(b + oQp1);
    p.qo[2] = (float *)(b + oQo2);
    p.qd[2] = (float *)(b + oQd2);
    p.qpath[2] = (uint32_t *)(b + oQp2);
    p.lastPdf = (float *)(b + oPdf);
    p.lastN = (float *)(b + oLastN);
    p.lastW = (float *)(b + oLastW);
    p.nee = (float *)(b + oNee);
    p.neeTarget = (uint32_t *)(b + oTgt);
    p.neeDist = (float *)(b + oDist);
    p.counters = (uint32_t *)(b + oCnt);
    p.metrics = (unsigned long long *)(b + oMet);
    p.ht = (float *)(b + oHt);
    p.hu = (float *)(b + oHu);
    p.hv = (float *)(b + oHv);
    p.hinst = (uint32_t *)(b + oHi);
    p.hprim = (uint32_t *)(b + oHp);
    bytes = off;
    return SPG_OK;
}

static int ensure_paths(spg_Device *d, uint32_t capacity, uint32_t bounces)
{
    if (d->pathsBlock && d->paths.capacity >= capacity &&
        d->paths.bounces >= bounces)
        return SPG_OK;
    free_paths(d);
    return alloc_path_state(d->paths, d->pathsBlock, d->pathsBytes, capacity,
                            bounces);
}

template <typename T>
static int ensure_buffer(void *&ptr, size_t &have, size_t count)
{
    if (ptr && have >= count) return SPG_OK;
    free_ptr(ptr);
    have = 0;
    HIPCHK(hipMalloc(&ptr, std::max<size_t>(count, 16) * sizeof(T)));
    have = count;
    return SPG_OK;
}

struct LaunchTimer
{
    spg_Device *d;
    int next = 0;
    int cls[(kEvents - 4) / 2];
    bool on = true;
    void begin(int c)
    {
        if (!on || next >= kEvents - 4) return;
        cls[next / 2] = c;
        hipEventRecord(d->ev[next], d->stream);
    }
    void end()
    {
        if (!on || next >= kEvents - 4) return;
        hipEventRecord(d->ev[next + 1], d->stream);
        next += 2;
    }
};

// devImage == nullptr: render into the device's staging image (sized here,
// under the device lock) and copy the tiles out to hostImage.
static int trace(const spg_Context *ctx, const spg_Tile *tiles, uint32_t count,
                 spg_Metrics *metrics, float *devImage, float *hostImage)
{
    if (!ctx || !ctx->camera || !ctx->scene || (!tiles && count))
        return fail(SPG_E_INVALID, "trace: NULL context, camera, scene or tiles");
    spg_Scene *s = ctx->scene;
    spg_Device *d = s->dev;
    const spg_Camera &c = *ctx->camera;
    const spg_TraceConfig &cfg = ctx->config;
    if (!s->built) return fail(SPG_E_STATE, "trace: scene not built");
    if (cfg.samplesPerPixel == 0 || cfg.bounces == 0 || cfg.bounces > kMaxBounces)
        return fail(SPG_E_INVALID, "trace: samplesPerPixel must be > 0 and bounces in 1..16");
    if (cfg.rngMode != SPG_RNG_XORSHIFT_PIXEL && cfg.rngMode != SPG_RNG_PCG)
        return fail(SPG_E_INVALID, "trace: unknown rngMode");
    if (cfg.estimator != SPG_EST_PARITY && cfg.estimator != SPG_EST_UNBIASED &&
        cfg.estimator != SPG_EST_GUIDED && cfg.estimator != SPG_EST_NEE &&
        cfg.estimator != SPG_EST_GUIDED_NEE)
        return fail(SPG_E_INVALID, "trace: unknown estimator");
    const bool guided =
        cfg.estimator == SPG_EST_GUIDED || cfg.estimator == SPG_EST_GUIDED_NEE;
    const bool nee = cfg.estimator == SPG_EST_NEE || cfg.estimator == SPG_EST_GUIDED_NEE;
    if (guided && (!cfg.guide || cfg.guide->dev != d))
        return fail(SPG_E_INVALID, "trace: the guided estimator needs a guide on the scene's device");
    if (cfg.transcendentals != SPG_TX_DOUBLE && cfg.transcendentals != SPG_TX_FAST)
        return fail(SPG_E_INVALID, "trace: unknown transcendentals mode");
    if (c.width == 0 || c.height == 0)
        return fail(SPG_E_INVALID, "trace: camera not configured");
    if (!devImage && !hostImage)
        return fail(SPG_E_INVALID, "trace: no output image");
    auto t0 = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    if (s->dirty) CHK(upload_tables(s));

    // pixel slots in tile order (clamped to the image as
    // sp_PathTraceTile clamps, src/simd_path_tracer.cpp:186-189); kept from
    // the previous call when the tiles and the image size are the same
    const bool sameTiles =
        d->slotW == c.width && d->slotH == c.height && d->slotTiles.size() == count &&
        (count == 0 || memcmp(d->slotTiles.data(), tiles, count * sizeof(spg_Tile)) == 0);
    if (!sameTiles)
    {
        d->slots.clear();
        for (uint32_t k = 0; k < count; ++k)
        {
            uint32_t x1 = std::min(tiles[k].maxX, c.width);
            uint32_t y1 = std::min(tiles[k].maxY, c.height);
            for (uint32_t y = tiles[k].minY; y < y1; ++y)
                for (uint32_t x = tiles[k].minX; x < x1; ++x)
                    d->slots.push_back(y * c.width + x);
        }
        d->slotTiles.assign(tiles, tiles + count);
        d->slotW = c.width;
        d->slotH = c.height;
        d->slotsOnDevice = false;
    }
    const std::vector<uint32_t> &slots = d->slots;
    if (!devImage)
    {
        // host pixels: the tiles' pixels gathered in slot order on the
        // device, then one copy into pinned host memory and a scatter of
        // the tile rows into the caller's image (a strided copy per tile
        // into pageable memory costs a runtime round trip per row)
        const size_t bytes = std::max<size_t>(slots.size(), 1) * 16;
        CHK(ensure_buffer<float>(d->image, d->imageCount, bytes / 4));
        devImage = (float *)d->image;
        if (d->hostStageBytes < bytes)
        {
            if (d->hostStage) hipHostFree(d->hostStage);
            d->hostStage = nullptr;
            d->hostStageBytes = 0;
            HIPCHK(hipHostMalloc(&d->hostStage, bytes, hipHostMallocDefault));
            d->hostStageBytes = bytes;
        }
    }
    const uint32_t spp = cfg.samplesPerPixel;
    const bool pcg = cfg.rngMode == SPG_RNG_PCG;
    uint64_t want = (uint64_t)slots.size() * (pcg ? spp : 1);
    uint32_t cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, 1),
                                                kMaxPathsInFlight);
    CHK(ensure_paths(d, cap, cfg.bounces));
    cap = d->paths.capacity;
    const size_t slotHave = d->slotPixelCount;
    CHK(ensure_buffer<uint32_t>(d->slotPixel, d->slotPixelCount, cap));
    if (d->slotPixelCount != slotHave) d->slotsOnDevice = false;  // reallocated
    CHK(ensure_buffer<float>(d->accum, d->accumCount, 3 * (size_t)cap));
    HIPCHK(launch_clear_counters(d->stream, (uint32_t *)d->paths.metrics, 32));

    TraceParams p{};
    p.scene = s->dev_scene;
    p.cam = dev_camera(c);
    p.paths = d->paths;
    p.slotPixel = (const uint32_t *)d->slotPixel;
    p.accum = (float *)d->accum;
    p.spp = spp;
    p.bounces = cfg.bounces;
    p.clampValue = cfg.radianceClamp;
    p.estimator = cfg.estimator;
    p.rngMode = cfg.rngMode;
    p.seed = cfg.seed;
    p.tx = cfg.transcendentals;
    p.collect = cfg.collectMetrics;
    p.work = d->persist.counter;
    const bool train = guided && cfg.guideTrain;
    if (guided)
        CHK(guide_prepare_trace(cfg.guide, s, cap, cfg.bounces, train, p.guide));
    if (cfg.estimator == SPG_EST_GUIDED_NEE)
    {
        // the guided tracer with next-event estimation learns the split of
        // its light samples between the sky and the emitters: each kind's
        // share of the light its training passes' samples delivered
        p.learnLight = train ? 1u : 0u;
        if (p.scene.envTex >= 0 && p.scene.emitCount > 0)
            p.scene.envSelect = guide_light_split(cfg.guide, p.scene.envSelect);
    }
    else if (cfg.estimator == SPG_EST_NEE && s->learnLight)
    {
        // the same for the NEE tracer, learned by the scene from its calls
        p.learnLight = 1u;
        if (p.scene.envTex >= 0 && p.scene.emitCount > 0)
            p.scene.envSelect = light_split(s->lightStat, p.scene.envSelect);
    }

    LaunchTimer tm{d};
    uint32_t launches[6] = {0};
    hipEventRecord(d->ev[kEvents - 2], d->stream);
    for (size_t s0 = 0; s0 < slots.size(); s0 += cap)
    {
        uint32_t nslots = (uint32_t)std::min<size_t>(cap, slots.size() - s0);
        const bool whole = s0 == 0 && slots.size() <= cap;
        if (!(whole && d->slotsOnDevice))
            HIPCHK(hipMemcpyAsync(d->slotPixel, slots.data() + s0, 4 * (size_t)nslots,
                                  hipMemcpyHostToDevice, d->stream));
        d->slotsOnDevice = whole;
        HIPCHK(launch_clear_floats(d->stream, p.accum, 3 * (size_t)nslots));
        p.slots = nslots;
        uint32_t sppPass = pcg ? std::max(1u, std::min(spp, cap / nslots)) : 1u;
        for (uint32_t done = 0; done < spp;)
        {
            uint32_t sp = std::min(sppPass, spp - done);
            p.sppPass = sp;
            p.sampleBase = cfg.sampleOffset + done;
            // accum and the paths are laid out with the slot count of this
            // chunk; accum rows are 3 x nslots
            if (guided)
                CHK(guide_begin_pass(cfg.guide, d->stream, p.guide,
                                     (p.sampleBase + 1u) * 0x9E3779B9u ^ cfg.seed ^
                                         (uint32_t)s0,
                                     nslots * sp));
            tm.begin(kRaygen);
            HIPCHK(launch_raygen(d->stream, p));
            tm.end();
            launches[kRaygen]++;
            int q = 0;
            for (uint32_t b = 0; b < cfg.bounces; ++b)
            {
                tm.begin(kExtend);
                HIPCHK(launch_extend(d->stream, p, q, d->persist));
                tm.end();
                // at the last bounce no hit scatters (a path at bounce b
                // holds b vertices), so there is nothing to infer
                if (guided && b + 1 < cfg.bounces)
                {
                    tm.begin(kGuide);
                    HIPCHK(launch_guide_infer(d->stream, p, q));
                    tm.end();
                    launches[kGuide]++;
                }
                tm.begin(kShade);
                HIPCHK(launch_shade(d->stream, p, q));
                tm.end();
                launches[kExtend]++;
                launches[kShade]++;
                if (nee && b + 1 < cfg.bounces)
                {
                    // shadow rays of this bounce, then their contributions
                    tm.begin(kExtend);
                    HIPCHK(launch_extend(d->stream, p, 2, d->persist));
                    tm.end();
                    tm.begin(kShade);
                    HIPCHK(launch_resolve_shadow(d->stream, p));
                    tm.end();
                    launches[kExtend]++;
                    launches[kShade]++;
                }
                q ^= 1;
            }
            tm.begin(kFinal);
            HIPCHK(launch_finalize(d->stream, p));
            tm.end();
            launches[kFinal]++;
            if (train)
            {
                tm.begin(kGuide);
                HIPCHK(launch_guide_collect(d->stream, p));
                CHK(guide_train_pass(cfg.guide, d->stream));
                tm.end();
                launches[kGuide]++;
            }
            done += sp;
        }
        if (hostImage) HIPCHK(launch_write_slots(d->stream, p, devImage + 4 * s0));
        else HIPCHK(launch_write_pixels(d->stream, p, devImage));
    }
    hipEventRecord(d->ev[kEvents - 1], d->stream);
    unsigned long long m[16] = {0};
    HIPCHK(hipMemcpyAsync(m, d->paths.metrics, 128, hipMemcpyDeviceToHost,
                          d->stream));
    if (hostImage && !slots.empty())
        HIPCHK(hipMemcpyAsync(d->hostStage, devImage, slots.size() * 16,
                              hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    if (p.learnLight)
    {
        uint64_t *stat = cfg.estimator == SPG_EST_GUIDED_NEE ? cfg.guide->lightStat
                                                             : s->lightStat;
        stat[0] += m[8];
        stat[1] += m[9];
    }
    if (hostImage)
    {
        // slot order is tile order, row by row within a tile, so each
        // tile row is one contiguous run of the staging copy
        const char *src = (const char *)d->hostStage;
        for (uint32_t k = 0; k < count; ++k)
        {
            uint32_t x1 = std::min(tiles[k].maxX, c.width);
            uint32_t y1 = std::min(tiles[k].maxY, c.height);
            if (x1 <= tiles[k].minX || y1 <= tiles[k].minY) continue;
            const size_t run = (size_t)(x1 - tiles[k].minX) * 16;
            for (uint32_t y = tiles[k].minY; y < y1; ++y, src += run)
                memcpy((char *)hostImage + ((size_t)y * c.width + tiles[k].minX) * 16, src,
                       run);
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    for (int k = 0; k < 6; ++k) d->lastMs[k] = 0.0;
    for (int k = 0; k < tm.next; k += 2)
    {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, d->ev[k], d->ev[k + 1]);
        d->lastMs[tm.cls[k / 2]] += ms;
    }
    float total = 0.0f;
    hipEventElapsedTime(&total, d->ev[kEvents - 2], d->ev[kEvents - 1]);
    d->lastMs[5] = total;
    for (int k = 0; k < 6; ++k) d->lastLaunches[k] = launches[k];
    for (int k = 0; k < 8; ++k) d->lastCounters[k] = m[k];
    if (metrics)
    {
        metrics->values[SPG_METRIC_ELAPSED_NS] += (uint64_t)
            std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
        metrics->values[SPG_METRIC_PATHS_TRACED] += (uint64_t)slots.size() * spp;
        metrics->values[SPG_METRIC_RAYS_TRACED] += m[0];
        metrics->values[SPG_METRIC_RAY_HIT_COUNT] += m[1];
        metrics->values[SPG_METRIC_RAY_MISS_COUNT] += m[2];
        metrics->values[SPG_METRIC_MIDPHASE_AABB_TESTS] += m[3];
        metrics->values[SPG_METRIC_MESH_TESTS] += m[4];
    }
    return SPG_OK;
}

// Whether two host-pixel requests can be traced as one wavefront: the
// same scene, camera, options and host image.  A guided call that trains
// changes the guide, so it always runs alone.
static bool combinable(const TileRequest &a, const TileRequest &b)
{
    const spg_TraceConfig &ca = a.ctx->config, &cb = b.ctx->config;
    if ((ca.estimator == SPG_EST_GUIDED || ca.estimator == SPG_EST_GUIDED_NEE) &&
        (ca.guideTrain || cb.guideTrain))
        return false;
    return a.ctx->scene == b.ctx->scene && a.hostPixels == b.hostPixels &&
           memcmp(a.ctx->camera, b.ctx->camera, sizeof(spg_Camera)) == 0 &&
           memcmp(&ca, &cb, sizeof(spg_TraceConfig)) == 0;
}

// Traces a group of combinable requests in one call and hands each its
// share of the counters: paths exactly (its pixels x spp), rays, hits,
// misses and test counts in proportion to its paths (the shares add up to
// the group's totals), elapsed time from its own start.
static void trace_group(spg_Device *d, std::vector<TileRequest *> &group)
{
    std::vector<spg_Tile> tiles;
    for (TileRequest *r : group) tiles.insert(tiles.end(), r->tiles, r->tiles + r->count);
    spg_Metrics m{};
    const TileRequest &lead = *group[0];
    int st = trace(lead.ctx, tiles.data(), (uint32_t)tiles.size(), &m, nullptr,
                   lead.hostPixels);
    std::string err = st ? std::string(spg_last_error()) : std::string();
    auto t1 = std::chrono::steady_clock::now();
    const spg_Camera &c = *lead.ctx->camera;
    const uint64_t spp = lead.ctx->config.samplesPerPixel;
    std::vector<uint64_t> paths(group.size());
    uint64_t allPaths = 0;
    for (size_t k = 0; k < group.size(); ++k)
    {
        uint64_t px = 0;
        for (uint32_t j = 0; j < group[k]->count; ++j)
        {
            const spg_Tile &t = group[k]->tiles[j];
            uint64_t x1 = std::min(t.maxX, c.width), y1 = std::min(t.maxY, c.height);
            if (x1 > t.minX && y1 > t.minY) px += (x1 - t.minX) * (y1 - t.minY);
        }
        paths[k] = px * spp;
        allPaths += paths[k];
    }
    const int shared[] = {SPG_METRIC_RAYS_TRACED, SPG_METRIC_RAY_HIT_COUNT,
                          SPG_METRIC_RAY_MISS_COUNT, SPG_METRIC_MIDPHASE_AABB_TESTS,
                          SPG_METRIC_MESH_TESTS};
    uint64_t given[5] = {0, 0, 0, 0, 0};
    uint64_t before = 0;
    std::lock_guard<std::mutex> g(d->qlock);
    for (size_t k = 0; k < group.size(); ++k)
    {
        TileRequest *r = group[k];
        if (!st && r->metrics)
        {
            spg_Metrics &o = *r->metrics;
            o.values[SPG_METRIC_ELAPSED_NS] += (uint64_t)
                std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - r->start).count();
            o.values[SPG_METRIC_PATHS_TRACED] += paths[k];
            before += paths[k];
            for (int j = 0; j < 5; ++j)
            {
                // cumulative share up to this request, so the shares sum
                // exactly to the group's count
                uint64_t tot = m.values[shared[j]];
                uint64_t upto = k + 1 == group.size() || allPaths == 0
                                    ? tot
                                    : (uint64_t)((unsigned __int128)tot * before / allPaths);
                o.values[shared[j]] += upto - given[j];
                given[j] = upto;
            }
        }
        r->status = st;
        r->error = err;
        r->done = true;
    }
}

// Host-pixel trace through the device's request queue: the caller that
// finds no trace running takes every queued request combinable with the
// oldest one, traces them as one wavefront and wakes their callers; the
// others wait.  Concurrent callers on one context (the reference's worker
// threads, src/main.cpp:731-759) therefore share launch sequences instead
// of queueing whole sequences behind the device lock.
static int host_trace(const spg_Context *ctx, const spg_Tile *tiles,
                      uint32_t count, spg_Metrics *metrics, float *hostPixels)
{
    if (!ctx || !ctx->camera || !ctx->scene || !hostPixels || (!tiles && count))
        return fail(SPG_E_INVALID, "trace: NULL argument");
    spg_Device *d = ctx->scene->dev;
    TileRequest req{ctx, tiles, count, metrics, hostPixels,
                    std::chrono::steady_clock::now()};
    std::unique_lock<std::mutex> q(d->qlock);
    d->queue.push_back(&req);
    d->qcv.notify_all();  // a leader gathering requests may be waiting
    while (!req.done)
    {
        if (d->leaderActive)
        {
            d->qcv.wait(q);
            continue;
        }
        d->leaderActive = true;
        // the callers of the last combined trace come back with their
        // next tiles within microseconds: wait briefly for as many
        // requests as it had, so a pool of callers keeps tracing its tiles
        // together instead of one or two at a time
        const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(500);
        while (d->queue.size() < d->lastGroup &&
               d->qcv.wait_until(q, until) != std::cv_status::timeout)
        {
        }
        std::vector<TileRequest *> group;
        TileRequest *first = d->queue.front();
        for (auto it = d->queue.begin(); it != d->queue.end();)
        {
            if (*it == first || combinable(*first, **it))
            {
                group.push_back(*it);
                it = d->queue.erase(it);
            }
            else
                ++it;
        }
        d->lastGroup = group.size();
        q.unlock();
        try
        {
            trace_group(d, group);
        }
        catch (...)
        {
            // host allocation failure: fail the group, keep the queue alive
            std::lock_guard<std::mutex> g(d->qlock);
            for (TileRequest *r : group)
                if (!r->done)
                {
                    r->status = SPG_E_CAPACITY;
                    r->error = "trace: host allocation failed";
                    r->done = true;
                }
        }
        q.lock();
        d->leaderActive = false;
        d->qcv.notify_all();
    }
    if (req.status) fail(req.status, req.error);
    return req.status;
}

}  // namespace spg

extern "C" {

int spg_path_trace_tile(const spg_Context *ctx, spg_Tile tile,
                        spg_Metrics *metrics, float *hostPixels)
{
    return host_trace(ctx, &tile, 1, metrics, hostPixels);
}

int spg_path_trace_tiles(const spg_Context *ctx, const spg_Tile *tiles,
                         uint32_t count, spg_Metrics *metrics,
                         float *hostPixels)
{
    return host_trace(ctx, tiles, count, metrics, hostPixels);
}

int spg_path_trace_tiles_device(const spg_Context *ctx, const spg_Tile *tiles,
                                uint32_t count, spg_Metrics *metrics,
                                float *devicePixels)
{
    if (!devicePixels) return fail(SPG_E_INVALID, "trace: NULL device image");
    return trace(ctx, tiles, count, metrics, devicePixels, nullptr);
}

int spg_last_trace_timings(spg_Device *d, double *ms6, uint32_t *launches6)
{
    if (!d) return fail(SPG_E_INVALID, "NULL device");
    std::lock_guard<std::mutex> g(d->lock);
    for (int k = 0; k < 6; ++k)
    {
        if (ms6) ms6[k] = d->lastMs[k];
        if (launches6) launches6[k] = d->lastLaunches[k];
    }
    return SPG_OK;
}

int spg_last_trace_counters(spg_Device *d, uint64_t *counters8)
{
    if (!d || !counters8) return fail(SPG_E_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(d->lock);
    for (int k = 0; k < 8; ++k) counters8[k] = d->lastCounters[k];
    return SPG_OK;
}

// ------------------------------------------------------------ closest hit
static int check_mesh(spg_Scene *s, uint32_t meshId)
{
    if (!s) return fail(SPG_E_INVALID, "NULL scene");
    if (!s->built) return fail(SPG_E_STATE, "scene not built");
    if (meshId >= s->meshes.size())
        return fail(SPG_E_NOT_FOUND, "unknown mesh id");
    return SPG_OK;
}

static DevMesh dev_mesh(const HostMesh &m)
{
    DevMesh d{};
    d.nodes = (const float4 *)m.dNodes;
    d.tris = (const float4 *)m.dTris;
    d.verts = (const float *)m.dVerts;
    d.idx = (const uint32_t *)m.dIdx;
    d.shade = (const float4 *)m.dShade;
    d.smooth = m.smooth;
    d.triCount = (uint32_t)(m.idx.size() / 3);
    return d;
}

int spg_intersect_closest_device(spg_Scene *s, uint32_t meshId,
                                 const float *ox, const float *oy,
                                 const float *oz, const float *dx,
                                 const float *dy, const float *dz, uint32_t n,
                                 uint32_t *primId, float *t, float *u, float *v,
                                 spg_TraversalStats *stats)
{
    CHK(check_mesh(s, meshId));
    if (n && (!ox || !oy || !oz || !dx || !dy || !dz || !primId || !t))
        return fail(SPG_E_INVALID, "spg_intersect_closest: NULL array");
    spg_Device *d = s->dev;
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    DevMesh m = dev_mesh(s->meshes[meshId]);
    unsigned long long *st = nullptr;
    if (stats)
    {
        CHK(ensure_paths(d, 1024, 1));
        st = d->paths.metrics;
        HIPCHK(launch_clear_counters(d->stream, (uint32_t *)st, 16));
    }
    HIPCHK(launch_closest_mesh(d->stream, &m, ox, oy, oz, dx, dy, dz, n, primId,
                               t, u, v, st, d->persist));
    if (stats)
    {
        unsigned long long h[8];
        HIPCHK(hipMemcpyAsync(h, st, 64, hipMemcpyDeviceToHost, d->stream));
        HIPCHK(hipStreamSynchronize(d->stream));
        stats->rays = h[0];
        stats->boxTests = h[1];
        stats->triTests = h[2];
        stats->nodeFetches = h[1] / 2;
        stats->hits = h[4];
    }
    return hip_check(hipStreamSynchronize(d->stream), "closest hit");
}

int spg_intersect_closest_device_noalloc(spg_Scene *s, uint32_t meshId,
                                         const float *ox, const float *oy,
                                         const float *oz, const float *dx,
                                         const float *dy, const float *dz,
                                         uint32_t n, uint32_t *primId, float *t,
                                         float *u, float *v)
{
    CHK(check_mesh(s, meshId));
    if (n && (!ox || !oy || !oz || !dx || !dy || !dz || !primId || !t))
        return fail(SPG_E_INVALID, "spg_intersect_closest_device_noalloc: NULL array");
    const HostMesh &hm = s->meshes[meshId];
    if (hm.widePending > kNoAllocStack)
        return fail(SPG_E_CAPACITY,
                    "spg_intersect_closest_device_noalloc: the mesh's walk needs " +
                        std::to_string(hm.widePending) + " stack entries, the static "
                        "scratch holds " + std::to_string(kNoAllocStack));
    spg_Device *d = s->dev;
    if (!d->noallocScratch)
        return fail(SPG_E_STATE, "spg_intersect_closest_device_noalloc: no scratch address");
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    DevMesh m = dev_mesh(hm);
    HIPCHK(launch_closest_mesh_noalloc(d->stream, &m, ox, oy, oz, dx, dy, dz, n, primId,
                                       t, u, v, d->noallocScratch));
    return hip_check(hipStreamSynchronize(d->stream), "closest hit");
}

int spg_intersect_closest(spg_Scene *s, uint32_t meshId, const float *ox,
                          const float *oy, const float *oz, const float *dx,
                          const float *dy, const float *dz, uint32_t n,
                          uint32_t *primId, float *t, float *u, float *v)
{
    CHK(check_mesh(s, meshId));
    if (n == 0) return SPG_OK;
    spg_Device *d = s->dev;
    void *buf = nullptr;
    {
        std::lock_guard<std::mutex> g(d->lock);
        HIPCHK(hipSetDevice(d->ordinal));
        HIPCHK(hipMalloc(&buf, (size_t)n * 4 * 10));
    }
    float *f = (float *)buf;
    const float *src[6] = {ox, oy, oz, dx, dy, dz};
    int r = SPG_OK;
    for (int k = 0; k < 6 && !r; ++k)
        r = hip_check(hipMemcpy(f + (size_t)k * n, src[k], 4 * (size_t)n,
                                hipMemcpyHostToDevice), "upload rays");
    uint32_t *dp = (uint32_t *)(f + 6 * (size_t)n);
    float *dt = f + 7 * (size_t)n, *du = f + 8 * (size_t)n, *dv = f + 9 * (size_t)n;
    if (!r)
        r = spg_intersect_closest_device(s, meshId, f, f + n, f + 2 * (size_t)n,
                                         f + 3 * (size_t)n, f + 4 * (size_t)n,
                                         f + 5 * (size_t)n, n, dp, dt, du, dv,
                                         nullptr);
    if (!r) r = hip_check(hipMemcpy(primId, dp, 4 * (size_t)n, hipMemcpyDeviceToHost), "d2h");
    if (!r) r = hip_check(hipMemcpy(t, dt, 4 * (size_t)n, hipMemcpyDeviceToHost), "d2h");
    if (!r && u) r = hip_check(hipMemcpy(u, du, 4 * (size_t)n, hipMemcpyDeviceToHost), "d2h");
    if (!r && v) r = hip_check(hipMemcpy(v, dv, 4 * (size_t)n, hipMemcpyDeviceToHost), "d2h");
    hipFree(buf);
    return r;
}

int spg_accumulate_device(spg_Device *d, float *dst, const float *src,
                          uint32_t pixels, uint32_t before, uint32_t added)
{
    if (!d || (pixels && (!dst || !src)) || before + (uint64_t)added == 0)
        return fail(SPG_E_INVALID, "spg_accumulate_device: bad argument");
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    HIPCHK(launch_accumulate(d->stream, (float4 *)dst, (const float4 *)src, pixels,
                             before, added));
    return hip_check(hipStreamSynchronize(d->stream), "accumulate");
}

int spg_tonemap_device(spg_Device *d, const float *rgba, uint32_t pixels,
                       float exposure, int tx, uint8_t *out)
{
    if (!d || (pixels && (!rgba || !out)))
        return fail(SPG_E_INVALID, "spg_tonemap_device: bad argument");
    if (tx != SPG_TX_DOUBLE && tx != SPG_TX_FAST)
        return fail(SPG_E_INVALID, "spg_tonemap_device: transcendentals");
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    HIPCHK(launch_tonemap(d->stream, (const float4 *)rgba, pixels, exposure,
                          tx == SPG_TX_DOUBLE, (uchar4 *)out));
    return hip_check(hipStreamSynchronize(d->stream), "tonemap");
}

int spg_time_post(spg_Device *d, float *dst, const float *src, uint32_t pixels,
                  float exposure, int tx, uint8_t *out, uint32_t iterations,
                  double *accMs, double *tmMs)
{
    if (!d || !pixels || !dst || !src || !out || !iterations || !accMs || !tmMs)
        return fail(SPG_E_INVALID, "spg_time_post: bad argument");
    if (tx != SPG_TX_DOUBLE && tx != SPG_TX_FAST)
        return fail(SPG_E_INVALID, "spg_time_post: transcendentals");
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    hipEvent_t e0 = d->ev[kEvents - 4], e1 = d->ev[kEvents - 3],
               e2 = d->ev[kEvents - 2];
    HIPCHK(hipEventRecord(e0, d->stream));
    for (uint32_t i = 0; i < iterations; ++i)
        HIPCHK(launch_accumulate(d->stream, (float4 *)dst, (const float4 *)src,
                                 pixels, i + 1, 1));
    HIPCHK(hipEventRecord(e1, d->stream));
    for (uint32_t i = 0; i < iterations; ++i)
        HIPCHK(launch_tonemap(d->stream, (const float4 *)dst, pixels, exposure,
                              tx == SPG_TX_DOUBLE, (uchar4 *)out));
    HIPCHK(hipEventRecord(e2, d->stream));
    HIPCHK(hipEventSynchronize(e2));
    float a = 0.0f, b = 0.0f;
    HIPCHK(hipEventElapsedTime(&a, e0, e1));
    HIPCHK(hipEventElapsedTime(&b, e1, e2));
    *accMs = a / iterations;
    *tmMs = b / iterations;
    return SPG_OK;
}

int spg_sincos_host(const float *x, uint32_t n, float *s, float *c)
{
    if (n && (!x || !s || !c)) return fail(SPG_E_INVALID, "spg_sincos_host: NULL array");
    for (uint32_t i = 0; i < n; ++i) tsincos(x[i], &s[i], &c[i]);
    return SPG_OK;
}

int spg_env_table_build(const float *rgba, uint32_t faceWidth, uint32_t faceHeight,
                        float *pdf, float *prob, uint32_t *alias)
{
    if (!rgba || !pdf || !prob || !alias || faceWidth < 2 || faceHeight < 2)
        return fail(SPG_E_INVALID, "spg_env_table_build: bad argument");
    std::vector<float> a, b;
    std::vector<uint32_t> c;
    if (!build_env_table(rgba, faceWidth, faceHeight, a, b, c))
        return fail(SPG_E_INVALID, "spg_env_table_build: no positive finite texel");
    memcpy(pdf, a.data(), 4 * a.size());
    memcpy(prob, b.data(), 4 * b.size());
    memcpy(alias, c.data(), 4 * c.size());
    return SPG_OK;
}

int spg_write_pfm(const char *path, uint32_t width, uint32_t height,
                  const float *rgba)
{
    if (!path || !rgba || width == 0 || height == 0)
        return fail(SPG_E_INVALID, "spg_write_pfm: bad argument");
    FILE *f = fopen(path, "wb");
    if (!f) return fail(SPG_E_INVALID, "spg_write_pfm: cannot open the file");
    fprintf(f, "PF\n%u %u\n-1.0\n", width, height);
    std::vector<float> row(3 * (size_t)width);
    bool ok = true;
    for (uint32_t y = height; y-- > 0 && ok;)
    {
        const float *src = rgba + 4 * (size_t)y * width;
        for (uint32_t x = 0; x < width; ++x)
        {
            row[3 * x] = src[4 * x];
            row[3 * x + 1] = src[4 * x + 1];
            row[3 * x + 2] = src[4 * x + 2];
        }
        ok = fwrite(row.data(), 4, row.size(), f) == row.size();
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? SPG_OK : fail(SPG_E_INVALID, "spg_write_pfm: write failed");
}

// CreateIrradianceCubeMap, src/cubemap.cpp:108-233 (uniform sampling).  The
// sample grid is the reference's float loop counters, generated here the
// same way and handed to the kernel.
int spg_create_irradiance_cubemap(spg_Device *d, const float *equirect,
                                  uint32_t ew, uint32_t eh, uint32_t fw,
                                  uint32_t fh, float radianceClamp, int tx,
                                  float *out, double *kernelMs)
{
    if (!d || !equirect || !out || ew == 0 || eh == 0 || fw == 0 || fh == 0)
        return fail(SPG_E_INVALID, "spg_create_irradiance_cubemap: bad argument");
    if ((uint64_t)ew * eh > (1ull << 28) || (uint64_t)fw * fh > (1ull << 24))
        return fail(SPG_E_CAPACITY, "spg_create_irradiance_cubemap: image too large");
    if (tx != SPG_TX_DOUBLE && tx != SPG_TX_FAST)
        return fail(SPG_E_INVALID, "spg_create_irradiance_cubemap: transcendentals");
    std::vector<float> grid;
    const float sampleDelta = 0.1f;
    for (float phi = 0.0f; phi < 2.0f * kPi; phi += sampleDelta) grid.push_back(phi);
    const uint32_t nPhi = (uint32_t)grid.size();
    for (float theta = 0.0f; theta < 0.5f * kPi; theta += sampleDelta)
        grid.push_back(theta);
    const uint32_t nTheta = (uint32_t)grid.size() - nPhi;
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    const size_t srcBytes = (size_t)ew * eh * 16, outBytes = (size_t)6 * fw * fh * 16;
    void *buf = nullptr;
    HIPCHK(hipMalloc(&buf, srcBytes + outBytes + 4 * grid.size()));
    char *b = (char *)buf;
    float4 *dsrc = (float4 *)b, *dout = (float4 *)(b + srcBytes);
    float *dgrid = (float *)(b + srcBytes + outBytes);
    int r = hip_check(hipMemcpyAsync(dsrc, equirect, srcBytes, hipMemcpyHostToDevice,
                                     d->stream), "upload equirect");
    if (!r)
        r = hip_check(hipMemcpyAsync(dgrid, grid.data(), 4 * grid.size(),
                                     hipMemcpyHostToDevice, d->stream), "upload grid");
    if (!r) r = hip_check(hipEventRecord(d->ev[0], d->stream), "event");
    if (!r)
        r = hip_check(launch_irradiance(d->stream, dsrc, ew, eh, fw, fh, dgrid, nPhi,
                                        dgrid + nPhi, nTheta, radianceClamp,
                                        tx == SPG_TX_DOUBLE, dout),
                      "irradiance kernel");
    if (!r) r = hip_check(hipEventRecord(d->ev[1], d->stream), "event");
    if (!r)
        r = hip_check(hipMemcpyAsync(out, dout, outBytes, hipMemcpyDeviceToHost,
                                     d->stream), "download");
    if (!r) r = hip_check(hipStreamSynchronize(d->stream), "irradiance");
    if (!r && kernelMs)
    {
        float ms = 0.0f;
        r = hip_check(hipEventElapsedTime(&ms, d->ev[0], d->ev[1]), "event time");
        *kernelMs = ms;
    }
    hipFree(buf);
    return r;
}

int spg_time_intersect_closest(spg_Scene *s, uint32_t meshId, const float *ox,
                               const float *oy, const float *oz,
                               const float *dx, const float *dy,
                               const float *dz, uint32_t n, uint32_t *primId,
                               float *t, uint32_t iterations,
                               double *meanKernelMs)
{
    CHK(check_mesh(s, meshId));
    if (iterations == 0 || !meanKernelMs)
        return fail(SPG_E_INVALID, "spg_time_intersect_closest: bad argument");
    spg_Device *d = s->dev;
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    DevMesh m = dev_mesh(s->meshes[meshId]);
    HIPCHK(hipEventRecord(d->ev[kEvents - 4], d->stream));
    for (uint32_t i = 0; i < iterations; ++i)
        HIPCHK(launch_closest_mesh(d->stream, &m, ox, oy, oz, dx, dy, dz, n,
                                   primId, t, nullptr, nullptr, nullptr,
                                   d->persist));
    HIPCHK(hipEventRecord(d->ev[kEvents - 3], d->stream));
    HIPCHK(hipEventSynchronize(d->ev[kEvents - 3]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, d->ev[kEvents - 4], d->ev[kEvents - 3]));
    *meanKernelMs = ms / iterations;
    return SPG_OK;
}

int spg_intersect_scene(spg_Scene *s, const float *rays6, uint32_t n,
                        float *out7, uint32_t *object, uint32_t *prim)
{
    if (!s || (n && (!rays6 || !out7)))
        return fail(SPG_E_INVALID, "spg_intersect_scene: NULL argument");
    if (!s->built) return fail(SPG_E_STATE, "scene not built");
    if (n == 0) return SPG_OK;
    spg_Device *d = s->dev;
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    if (s->dirty) CHK(upload_tables(s));
    void *buf = nullptr;
    HIPCHK(hipMalloc(&buf, (size_t)n * 4 * 15));
    float *dr = (float *)buf, *dout = dr + 6 * (size_t)n;
    uint32_t *dobj = (uint32_t *)(dout + 7 * (size_t)n), *dprim = dobj + n;
    int r = hip_check(hipMemcpy(dr, rays6, 24 * (size_t)n, hipMemcpyHostToDevice), "h2d");
    if (!r) r = hip_check(launch_intersect_scene(d->stream, s->dev_scene, dr, n,
                                                 dout, dobj, dprim, d->persist),
                                  "launch");
    if (!r) r = hip_check(hipStreamSynchronize(d->stream), "sync");
    if (!r) r = hip_check(hipMemcpy(out7, dout, 28 * (size_t)n, hipMemcpyDeviceToHost), "d2h");
    if (!r && object) r = hip_check(hipMemcpy(object, dobj, 4 * (size_t)n, hipMemcpyDeviceToHost), "d2h");
    if (!r && prim) r = hip_check(hipMemcpy(prim, dprim, 4 * (size_t)n, hipMemcpyDeviceToHost), "d2h");
    hipFree(buf);
    return r;
}

}  // extern "C"
