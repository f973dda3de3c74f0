// Sanitizer driver for the library's host code (built by the Makefile beside
// it with AddressSanitizer + UndefinedBehaviorSanitizer, or with
// ThreadSanitizer).  No GPU.  Exits 0 when every case ran; a sanitizer
// finding aborts the process with its report.
//
//   host_fuzz obj   the OBJ parser (spg_obj_parse) on exactly sized heap
//                   buffers: a valid file cut at every byte, byte
//                   substitutions with digits, signs, slashes, spaces and
//                   newlines, and seeded random text
//   host_fuzz bvh   the threaded SAH build and the four-wide collapse on
//                   random, degenerate and coincident triangle sets
//                   ("bvh quick": up to 30000 triangles)
//   host_fuzz env   the environment and emitter importance tables on
//                   random, zero, huge and sparse inputs
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "spg_bvh.h"
#include "spg_env_table.h"
#include "spg_error.h"

static int parse(const std::string &text)
{
    // an exactly sized copy: the byte after the text is outside the buffer
    char *buf = new char[text.size() ? text.size() : 1];
    memcpy(buf, text.data(), text.size());
    uint32_t nv = 0, ni = 0;
    int r = spg_obj_parse(buf, text.size(), nullptr, 0, nullptr, 0, &nv, &ni);
    if (r == SPG_OK)
    {
        std::vector<float> v((size_t)nv * 8 + 1);
        std::vector<uint32_t> i((size_t)ni + 1);
        r = spg_obj_parse(buf, text.size(), v.data(), nv, i.data(), ni, &nv, &ni);
        for (uint32_t k = 0; k < ni; ++k)
            if (i[k] >= nv) std::abort();  // an index past the vertices
    }
    delete[] buf;
    return r;
}

static int fuzz_obj()
{
    const std::string good =
        "# cube corner\no a\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.5 1e-3\n"
        "vn 0 0 1\nvn 0 0 -1\nvt 0.5 0.5\nf 1/1/1 2/1/1 3/1/1 4/1/1\n"
        "f 1//2 5//2 2//2\nf -1//-1 -2//-1 -3//-1\no b\nv 9 9 9\nf 1//1 2//1 6//1\n";
    if (parse(good) != SPG_OK) return 1;
    int cases = 0;
    for (size_t cut = 0; cut <= good.size(); ++cut, ++cases) parse(good.substr(0, cut));
    const char subst[] = "0123456789-+/ \n.eE";
    for (size_t at = 0; at < good.size(); ++at)
        for (const char *c = subst; *c; ++c, ++cases)
        {
            std::string t = good;
            t[at] = *c;
            parse(t);
            parse(t.substr(0, at + 1));  // ends on the substituted byte
        }
    std::mt19937 rng(12345);
    const char alphabet[] = "vnft/ -0123456789.\n";
    for (int k = 0; k < 20000; ++k, ++cases)
    {
        std::string t;
        int len = (int)(rng() % 200);
        for (int j = 0; j < len; ++j) t += alphabet[rng() % (sizeof alphabet - 1)];
        parse(t);
    }
    printf("obj: %d cases\n", cases);
    return 0;
}

static bool build(const std::vector<float> &tris, uint32_t threads, uint32_t leaf)
{
    const uint32_t n = (uint32_t)(tris.size() / 9);
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a)
        {
            float x = tris[9 * i + a], y = tris[9 * i + 3 + a], z = tris[9 * i + 6 + a];
            lo[3 * i + a] = std::fmin(x, std::fmin(y, z));
            hi[3 * i + a] = std::fmax(x, std::fmax(y, z));
        }
    spg::BvhBuildInput in{lo.data(), hi.data(), n, leaf, 16, threads, 42};
    spg::BvhBuildResult out;
    if (!spg::build_bvh(in, out)) return n == 0;
    std::vector<spg::WideNode> wide;
    uint32_t pending = 0;
    if (!spg::collapse_wide(out, wide, pending)) return false;
    // every primitive referenced exactly once
    std::vector<int> seen(n, 0);
    for (uint32_t p : out.prims)
    {
        if (p >= n || seen[p]++) return false;
    }
    return out.prims.size() == n;
}

static int fuzz_bvh(uint32_t maxN)
{
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    int cases = 0;
    const uint32_t sizes[] = {1, 2, 3, 7, 64, 1000, 30000, 120000};
    for (uint32_t n : sizes)
        if (n <= maxN)
        for (int kind = 0; kind < 4; ++kind, ++cases)
        {
            std::vector<float> t(9 * (size_t)n);
            for (uint32_t i = 0; i < n; ++i)
                for (int j = 0; j < 9; ++j)
                {
                    float v = U(rng);
                    if (kind == 1) v = (j % 3 == 0) ? v : 0.25f;       // flat, degenerate
                    if (kind == 2) v = 0.5f;                            // all coincident
                    if (kind == 3) v = v * 1e6f + (float)(i % 3) * 1e-3f; // wide spread
                    t[9 * i + j] = v;
                }
            for (uint32_t threads : {1u, 8u})
                if (!build(t, threads, 1 + (uint32_t)(rng() % 8)))
                {
                    fprintf(stderr, "bvh: invariant failed n=%u kind=%d\n", n, kind);
                    return 1;
                }
        }
    printf("bvh: %d cases\n", cases);
    return 0;
}

static int fuzz_env()
{
    std::mt19937 rng(3);
    int cases = 0;
    for (uint32_t size : {1u, 2u, 5u, 32u})
        for (int kind = 0; kind < 4; ++kind, ++cases)
        {
            std::vector<float> rgba(6 * (size_t)size * size * 4);
            for (size_t i = 0; i < rgba.size(); ++i)
            {
                float u = (float)(rng() % 1000) / 1000.0f;
                rgba[i] = kind == 0 ? u : kind == 1 ? 0.0f : kind == 2 ? u * 3e38f
                                                                       : (i % 97 ? 0.0f : u);
            }
            std::vector<float> pdf, prob;
            std::vector<uint32_t> alias;
            bool ok = spg::build_env_table(rgba.data(), size, size, pdf, prob, alias);
            if (ok)
                for (uint32_t a : alias)
                    if (a >= 6 * size * size) return 1;
        }
    // emitter table over a few instances of a quad mesh
    const float verts[] = {-1, 0, -1, 0, 1, 0, 0, 0, 1, 0, -1, 0, 1, 0, 1, 0,
                           1,  0, 1,  0, 1, 0, 1, 1, -1, 0, 1, 0, 1, 0, 0, 1};
    const uint32_t idx[] = {0, 1, 2, 2, 3, 0};
    float model[16] = {2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0, 1, 2, 3, 1};
    for (int k = 0; k < 50; ++k, ++cases)
    {
        std::vector<spg::EmitterInput> in;
        for (int j = 0; j < 1 + k % 5; ++j)
            in.push_back({model, verts, idx, 2, (rng() % 3) ? (double)(rng() % 10) : 0.0});
        std::vector<float> pdfA, prob;
        std::vector<uint32_t> alias, pair;
        std::vector<int32_t> base;
        spg::build_emitter_table(in, pdfA, prob, alias, pair, base);
        for (uint32_t a : alias)
            if (a >= pdfA.size()) return 1;
    }
    printf("env: %d cases\n", cases);
    return 0;
}

int main(int argc, char **argv)
{
    std::string what = argc > 1 ? argv[1] : "all";
    // "quick": triangle sets up to 30000 (for the slower ThreadSanitizer run)
    const uint32_t maxN = argc > 2 && std::string(argv[2]) == "quick" ? 30000u : 120000u;
    int r = 0;
    if (what == "obj" || what == "all") r |= fuzz_obj();
    if (what == "bvh" || what == "all") r |= fuzz_bvh(maxN);
    if (what == "env" || what == "all") r |= fuzz_env();
    return r;
}
