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

#include <vector>

namespace spg {

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
    prob.assign(n, 1.0f);
    alias.resize(n);
    std::vector<double> q(n);
    std::vector<uint32_t> small, large;
    for (size_t i = 0; i < n; ++i)
    {
        const double P = wt[i] / total;
        pdf[i] = (float)(P / area[i]);
        q[i] = P * (double)n;
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
    return true;
}

}  // namespace spg
