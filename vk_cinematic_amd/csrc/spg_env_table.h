// Importance-sampling table of a cube-map environment (host side).
//
// The cube lookup (sample_cube, the inverse of CreateCubeMap's texel map,
// src/cubemap.cpp:237-291) gives texel (ix, iy) of a face the face-plane
// rectangle fx in [(2 ix - 1)/w - 1, (2 ix + 1)/w - 1), fy in
// (1 - (2 iy + 1)/h, 1 - (2 iy - 1)/h], clipped to [-1, 1] (the first and
// last texels of a row or column are half and one and a half wide).  A
// texel is chosen with probability P_i proportional to its mean channel
// value times its rectangle's area over (1 + fx^2 + fy^2)^(3/2) at the
// rectangle's centre (its solid angle to first order), then a point is
// drawn uniformly in the rectangle; the direction's density per steradian
// is P_i / area_i * (1 + fx^2 + fy^2)^(3/2).
//
// Built in double in texel order, so the table is a deterministic function
// of the cube map: pdf[i] = P_i / area_i (float) and Vose's alias table
// (prob[i] float, alias[i]); small and large work lists are stacks filled
// in ascending texel order.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include <vector>

namespace spg {

// Vose's alias table of probabilities P (sum 1), n entries: prob[i] (float)
// and alias[i]; small and large work lists are stacks filled in ascending
// order, leftovers keep probability 1.
inline void build_alias(const std::vector<double> &P, std::vector<float> &prob,
                        std::vector<uint32_t> &alias)
{
    const size_t n = P.size();
    prob.assign(n, 1.0f);
    alias.resize(n);
    std::vector<double> q(n);
    std::vector<uint32_t> small, large;
    for (size_t i = 0; i < n; ++i)
    {
        q[i] = P[i] * (double)n;
        alias[i] = (uint32_t)i;
        if (q[i] < 1.0) small.push_back((uint32_t)i);
        else large.push_back((uint32_t)i);
    }
    while (!small.empty() && !large.empty())
    {
        const uint32_t s = small.back(), l = large.back();
        small.pop_back();
        large.pop_back();
        prob[s] = (float)q[s];
        alias[s] = l;
        q[l] = (q[l] + q[s]) - 1.0;
        if (q[l] < 1.0) small.push_back(l);
        else large.push_back(l);
    }
}

// Face-plane rectangle of texel (ix, iy) of a w x h face.
inline void env_texel_rect(uint32_t ix, uint32_t iy, uint32_t w, uint32_t h,
                           double &x0, double &x1, double &y0, double &y1)
{
    x0 = ix == 0 ? -1.0 : (2.0 * ix - 1.0) / w - 1.0;
    x1 = ix + 1 == w ? 1.0 : (2.0 * ix + 1.0) / w - 1.0;
    y1 = iy == 0 ? 1.0 : 1.0 - (2.0 * iy - 1.0) / h;
    y0 = iy + 1 == h ? -1.0 : 1.0 - (2.0 * iy + 1.0) / h;
}

// Returns false (and leaves the outputs empty) when the map has no
// positive, finite weight: the environment is then not sampled.
inline bool build_env_table(const float *rgba, uint32_t w, uint32_t h,
                            std::vector<float> &pdf, std::vector<float> &prob,
                            std::vector<uint32_t> &alias)
{
    pdf.clear();
    prob.clear();
    alias.clear();
    if (!rgba || w < 2 || h < 2) return false;
    const size_t n = 6 * (size_t)w * h;
    std::vector<double> wt(n), area(n);
    double total = 0.0;
    for (size_t i = 0; i < n; ++i)
    {
        const uint32_t ix = (uint32_t)(i % w), iy = (uint32_t)((i / w) % h);
        double x0, x1, y0, y1;
        env_texel_rect(ix, iy, w, h, x0, x1, y0, y1);
        const double a = (x1 - x0) * (y1 - y0);
        const double xc = 0.5 * (x0 + x1), yc = 0.5 * (y0 + y1);
        const double t = 1.0 + xc * xc + yc * yc;
        double lum = ((double)rgba[4 * i] + (double)rgba[4 * i + 1] +
                      (double)rgba[4 * i + 2]) / 3.0;
        if (!(lum > 0.0) || !std::isfinite(lum)) lum = 0.0;
        area[i] = a;
        wt[i] = lum * a / (t * sqrt(t));
        total += wt[i];
    }
    if (!(total > 0.0) || !std::isfinite(total)) return false;
    pdf.resize(n);
    std::vector<double> P(n);
    for (size_t i = 0; i < n; ++i)
    {
        P[i] = wt[i] / total;
        pdf[i] = (float)(P[i] / area[i]);
    }
    build_alias(P, prob, alias);
    return true;
}

// Emissive triangles (next-event estimation toward emitting geometry): for
// every instance whose material has a constant emission with a positive
// mean channel value, every triangle of its mesh in index order, weight =
// that mean x the triangle's world-space area (vertices transformed by the
// instance's column-major model matrix in double).  P_g = weight / sum
// (double, instance-then-triangle order); pdfA[g] = P_g / area_g (density
// per unit area; 0 for a degenerate triangle); alias table as above;
// pair[g] = (instance, triangle); base[i] = the first entry of instance i
// or -1.
struct EmitterInput
{
    const float *model;     // 16 floats, column-major
    const float *verts;     // 8 floats per vertex
    const uint32_t *idx;
    uint32_t triCount;
    double lum;             // mean emission channel, <= 0: not an emitter
};

inline bool build_emitter_table(const std::vector<EmitterInput> &inst,
                                std::vector<float> &pdfA, std::vector<float> &prob,
                                std::vector<uint32_t> &alias,
                                std::vector<uint32_t> &pair,
                                std::vector<int32_t> &base)
{
    pdfA.clear();
    prob.clear();
    alias.clear();
    pair.clear();
    base.assign(inst.size(), -1);
    std::vector<double> wt, area;
    for (size_t i = 0; i < inst.size(); ++i)
    {
        const EmitterInput &e = inst[i];
        if (!(e.lum > 0.0) || !std::isfinite(e.lum) || e.triCount == 0) continue;
        base[i] = (int32_t)wt.size();
        const float *m = e.model;
        for (uint32_t k = 0; k < e.triCount; ++k)
        {
            double w[3][3];
            for (int v = 0; v < 3; ++v)
            {
                const float *p = e.verts + 8 * (size_t)e.idx[3 * k + v];
                for (int r = 0; r < 3; ++r)
                    w[v][r] = (double)m[r] * p[0] + (double)m[4 + r] * p[1] +
                              (double)m[8 + r] * p[2] + (double)m[12 + r];
            }
            const double ux = w[1][0] - w[0][0], uy = w[1][1] - w[0][1],
                         uz = w[1][2] - w[0][2];
            const double vx = w[2][0] - w[0][0], vy = w[2][1] - w[0][1],
                         vz = w[2][2] - w[0][2];
            const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz,
                         cz = ux * vy - uy * vx;
            const double a = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            area.push_back(a);
            wt.push_back(std::isfinite(a) ? e.lum * a : 0.0);
            pair.push_back((uint32_t)i);
            pair.push_back(k);
        }
    }
    double total = 0.0;
    for (double w : wt) total += w;
    if (!(total > 0.0) || !std::isfinite(total))
    {
        pair.clear();
        base.assign(inst.size(), -1);
        return false;
    }
    const size_t n = wt.size();
    pdfA.resize(n);
    std::vector<double> P(n);
    for (size_t g = 0; g < n; ++g)
    {
        P[g] = wt[g] / total;
        pdfA[g] = area[g] > 0.0 ? (float)(P[g] / area[g]) : 0.0f;
    }
    build_alias(P, prob, alias);
    return true;
}

// The spatial emitter choice of next-event estimation: per emitter
// instance a bounding sphere and its power, so a shading point x can weigh
// the instances by power / max(d^2, r^2) (0 for one wholly behind its
// tangent plane) and then take a triangle of the chosen instance by area.
// rec: 8 floats per emitter instance (centre xyz, radius, power = mean
// emission x world area, 1 / area, first table entry and triangle count as
// u32 bits); slot: per scene instance its emitter index or -1; probL /
// aliasL: per table entry (the order of build_emitter_table) an alias table
// over its instance's triangles by area, local indices.  Instances with
// zero total area get no record; more than kMaxSpatialEmitters records
// leave rec empty (the spatial choice off).  `base` is
// build_emitter_table's.
constexpr uint32_t kMaxSpatialEmitters = 256;
inline void build_emitter_instances(const std::vector<EmitterInput> &inst,
                                    const std::vector<int32_t> &base,
                                    std::vector<float> &rec, std::vector<int32_t> &slot,
                                    std::vector<float> &probL,
                                    std::vector<uint32_t> &aliasL)
{
    rec.clear();
    slot.assign(inst.size(), -1);
    size_t entries = 0;
    for (size_t i = 0; i < inst.size(); ++i)
        if (base[i] >= 0) entries = std::max(entries, (size_t)base[i] + inst[i].triCount);
    probL.assign(entries, 1.0f);
    aliasL.assign(entries, 0u);
    for (size_t i = 0; i < inst.size(); ++i)
    {
        if (base[i] < 0) continue;
        const EmitterInput &e = inst[i];
        const float *m = e.model;
        std::vector<double> area(e.triCount);
        double total = 0.0, lo[3] = {INFINITY, INFINITY, INFINITY},
               hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t k = 0; k < e.triCount; ++k)
        {
            double w[3][3];
            for (int v = 0; v < 3; ++v)
            {
                const float *p = e.verts + 8 * (size_t)e.idx[3 * k + v];
                for (int r = 0; r < 3; ++r)
                {
                    w[v][r] = (double)m[r] * p[0] + (double)m[4 + r] * p[1] +
                              (double)m[8 + r] * p[2] + (double)m[12 + r];
                    lo[r] = std::min(lo[r], w[v][r]);
                    hi[r] = std::max(hi[r], w[v][r]);
                }
            }
            const double ux = w[1][0] - w[0][0], uy = w[1][1] - w[0][1],
                         uz = w[1][2] - w[0][2];
            const double vx = w[2][0] - w[0][0], vy = w[2][1] - w[0][1],
                         vz = w[2][2] - w[0][2];
            const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz,
                         cz = ux * vy - uy * vx;
            const double a = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            area[k] = std::isfinite(a) ? a : 0.0;
            total += area[k];
        }
        if (!(total > 0.0) || !std::isfinite(total)) continue;
        std::vector<double> P(e.triCount);
        for (uint32_t k = 0; k < e.triCount; ++k) P[k] = area[k] / total;
        std::vector<float> pr;
        std::vector<uint32_t> al;
        build_alias(P, pr, al);
        for (uint32_t k = 0; k < e.triCount; ++k)
        {
            probL[(size_t)base[i] + k] = pr[k];
            aliasL[(size_t)base[i] + k] = al[k];
        }
        const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        uint32_t first = (uint32_t)base[i], count = e.triCount;
        float fb, cb;
        memcpy(&fb, &first, 4);
        memcpy(&cb, &count, 4);
        slot[i] = (int32_t)(rec.size() / 8);
        rec.insert(rec.end(), {(float)(0.5 * (lo[0] + hi[0])), (float)(0.5 * (lo[1] + hi[1])),
                               (float)(0.5 * (lo[2] + hi[2])),
                               (float)(0.5 * sqrt(dx * dx + dy * dy + dz * dz)),
                               (float)(e.lum * total), (float)(1.0 / total), fb, cb});
    }
    if (rec.size() / 8 > kMaxSpatialEmitters)
    {
        // too many for a per-point scan: the spatial choice is off
        rec.clear();
        slot.assign(inst.size(), -1);
    }
}

}  // namespace spg
