// Wavefront OBJ ingestion for LoadMesh (src/mesh.cpp:5-62).  The reference
// imports through assimp with aiProcess_Triangulate and
// aiProcess_JoinIdenticalVertices and keeps positions and normals of the
// first mesh (texture coordinates are not copied, so they stay 0).  assimp
// is not part of this build; this parser produces the same triangles:
// polygons split into fans (0, k, k+1), one vertex per distinct (position,
// normal) pair in order of first use, the faces of the first object only.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "spg_error.h"

using namespace spg;

namespace {

struct ObjMesh
{
    std::vector<float> verts;    // 8 per vertex
    std::vector<uint32_t> idx;
};

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// One 1-based (or negative, relative) OBJ index; false if malformed.
bool obj_index(const char *&p, const char *end, long count, long &out)
{
    char *e = nullptr;
    long v = strtol(p, &e, 10);
    if (e == p) return false;
    p = e;
    if (v > 0) out = v - 1;
    else if (v < 0) out = count + v;
    else return false;
    return out >= 0 && out < count && p <= end;
}

int parse(const char *text, size_t len, ObjMesh &m)
{
    std::vector<float> pos, nrm;
    std::map<std::pair<long, long>, uint32_t> joined;
    int objects = 0;
    const char *p = text, *end = text + len;
    std::vector<uint32_t> poly;
    // each line is parsed from a NUL-terminated copy, so strtof/strtol stop
    // at the line's end even when the text is not NUL-terminated (e.g. an
    // exactly sized mapping whose last line ends in a digit)
    std::string line;
    while (p < end)
    {
        const char *nl = (const char *)memchr(p, '\n', (size_t)(end - p));
        if (!nl) nl = end;
        line.assign(p, (size_t)(nl - p));
        const char *q = line.c_str();
        const char *eol = q + line.size();
        while (q < eol && is_space(*q)) ++q;
        if (q + 1 < eol && q[0] == 'o' && is_space(q[1]))
        {
            if (++objects > 1) break;  // the first object only
        }
        else if (q + 2 < eol && q[0] == 'v' && (q[1] == ' ' || q[1] == '\t' ||
                                                (q[1] == 'n' && is_space(q[2]))))
        {
            const bool isN = q[1] == 'n';
            q += isN ? 2 : 1;
            float xyz[3];
            for (int k = 0; k < 3; ++k)
            {
                char *e = nullptr;
                xyz[k] = strtof(q, &e);
                if (e == q || e > eol)
                    return fail(SPG_E_INVALID, "spg_obj_parse: malformed vertex line");
                q = e;
            }
            (isN ? nrm : pos).insert((isN ? nrm : pos).end(), xyz, xyz + 3);
        }
        else if (q + 1 < eol && q[0] == 'f' && is_space(q[1]))
        {
            q += 1;
            poly.clear();
            const long np = (long)pos.size() / 3, nn = (long)nrm.size() / 3;
            for (;;)
            {
                while (q < eol && is_space(*q)) ++q;
                if (q >= eol) break;
                long vi = 0, ni = -1, ti = 0;
                if (!obj_index(q, eol, np, vi))
                    return fail(SPG_E_INVALID, "spg_obj_parse: bad position index");
                if (q < eol && *q == '/')
                {
                    ++q;
                    if (q < eol && *q != '/')
                    {
                        // a texture coordinate: checked, not kept
                        char *e = nullptr;
                        ti = strtol(q, &e, 10);
                        if (e == q || ti == 0)
                            return fail(SPG_E_INVALID, "spg_obj_parse: bad texcoord index");
                        q = e;
                    }
                    if (q < eol && *q == '/')
                    {
                        ++q;
                        if (!obj_index(q, eol, nn, ni))
                            return fail(SPG_E_INVALID, "spg_obj_parse: bad normal index");
                    }
                }
                if (ni < 0)
                    return fail(SPG_E_INVALID,
                                "spg_obj_parse: face corner without a normal "
                                "(LoadMesh requires normals)");
                auto key = std::make_pair(vi, ni);
                auto it = joined.find(key);
                uint32_t id;
                if (it == joined.end())
                {
                    id = (uint32_t)(m.verts.size() / 8);
                    joined.emplace(key, id);
                    const float v8[8] = {pos[3 * vi], pos[3 * vi + 1], pos[3 * vi + 2],
                                         nrm[3 * ni], nrm[3 * ni + 1], nrm[3 * ni + 2],
                                         0.0f, 0.0f};
                    m.verts.insert(m.verts.end(), v8, v8 + 8);
                }
                else
                    id = it->second;
                poly.push_back(id);
                if (q < eol && !is_space(*q))
                    return fail(SPG_E_INVALID, "spg_obj_parse: malformed face corner");
            }
            if (poly.size() < 3)
                return fail(SPG_E_INVALID, "spg_obj_parse: face with fewer than 3 corners");
            for (size_t k = 1; k + 1 < poly.size(); ++k)
            {
                m.idx.push_back(poly[0]);
                m.idx.push_back(poly[k]);
                m.idx.push_back(poly[k + 1]);
            }
        }
        p = nl < end ? nl + 1 : end;
    }
    if (m.idx.empty()) return fail(SPG_E_INVALID, "spg_obj_parse: no faces");
    return SPG_OK;
}

}  // namespace

int spg_obj_parse(const char *text, size_t len, float *verts8, uint32_t vertCap,
                  uint32_t *idx, uint32_t idxCap, uint32_t *nv, uint32_t *ni)
{
    if (!text || !nv || !ni) return fail(SPG_E_INVALID, "spg_obj_parse: NULL argument");
    ObjMesh m;
    CHK(parse(text, len, m));
    *nv = (uint32_t)(m.verts.size() / 8);
    *ni = (uint32_t)m.idx.size();
    if (!verts8 && !idx) return SPG_OK;  // sizes only
    if (!verts8 || !idx || vertCap < *nv || idxCap < *ni)
        return fail(SPG_E_CAPACITY, "spg_obj_parse: arrays too small");
    memcpy(verts8, m.verts.data(), m.verts.size() * sizeof(float));
    memcpy(idx, m.idx.data(), m.idx.size() * sizeof(uint32_t));
    return SPG_OK;
}
