// gm_merge.cpp -- the pairwise merging of superpixels into coarser levels
// (downsampleOps, gridoperators.h:59-138).  Host code (no device work), part
// of libgpumatting.so beside the device-side superpixel operators.
#include <stddef.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "gpumatting.h"

namespace gm {
gm_status fail(gm_status s, const char* fmt, ...);
}
using gm::fail;

extern "C" {

// downsampleOps, gridoperators.h:59-138, with what it leaves open filled in:
// adjacency (4-neighbour pixel edges) and the means of each new level.
// Float arithmetic as there: means accumulate in pixel order and divide by the
// count; the squared distance adds the three channel squares in order.
gm_status gm_sp_merge_levels(int* labels, int* nlabels, const float* im, int nch, size_t pitch, int W, int H,
                             int nlevels, float threshold) {
    if (!labels || !nlabels || !im || nch < 3 || W <= 0 || H <= 0 || pitch < (size_t)W || *nlabels <= 0 || nlevels < 1)
        return fail(GM_ERR_INVALID, "gm_sp_merge_levels: arguments");
    const long N = (long)W * H;
    int nl = *nlabels;
    for (long u = 0; u < N; ++u)
        if (labels[u] < 0 || labels[u] >= nl) return fail(GM_ERR_INVALID, "gm_sp_merge_levels: label out of range");
    std::vector<float> means;
    std::vector<int> size, map;
    std::vector<char> merged;
    for (int level = 1; level < nlevels; ++level) {
        means.assign(3 * (size_t)nl, 0.f);
        size.assign(nl, 0);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int c = labels[(long)y * W + x];
                const float* px = im + ((size_t)y * pitch + x) * nch;
                means[3 * c + 0] += px[0];
                means[3 * c + 1] += px[1];
                means[3 * c + 2] += px[2];
                ++size[c];
            }
        for (int c = 0; c < nl; ++c)
            if (size[c])
                for (int q = 0; q < 3; ++q) means[3 * c + q] /= (float)size[c];
        // adjacent pairs (c < d)
        std::vector<std::pair<int, int>> pairs;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int c = labels[(long)y * W + x];
                if (x + 1 < W) {
                    const int d = labels[(long)y * W + x + 1];
                    if (d != c) pairs.emplace_back(std::min(c, d), std::max(c, d));
                }
                if (y + 1 < H) {
                    const int d = labels[(long)(y + 1) * W + x];
                    if (d != c) pairs.emplace_back(std::min(c, d), std::max(c, d));
                }
            }
        std::sort(pairs.begin(), pairs.end());
        pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
        struct Cand {
            float d;
            int u, v;
        };
        std::vector<Cand> cand;
        cand.reserve(pairs.size());
        for (const auto& pr : pairs) {
            const int u = pr.first, v = pr.second;
            float t = means[3 * u + 0] - means[3 * v + 0];
            t *= t;
            float dd = t;
            t = means[3 * u + 1] - means[3 * v + 1];
            t *= t;
            dd += t;
            t = means[3 * u + 2] - means[3 * v + 2];
            t *= t;
            dd += t;
            cand.push_back({dd, u, v});
        }
        // the reference's repeated scan for the first minimum over unmerged
        // pairs in (u, v) order is this ascending order
        std::stable_sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
            if (a.d != b.d) return a.d < b.d;
            return a.u != b.u ? a.u < b.u : a.v < b.v;
        });
        merged.assign(nl, 0);
        map.assign(nl, -1);
        int cur = 0;
        for (const Cand& c : cand) {
            if (!(c.d <= threshold)) break;
            if (merged[c.u] || merged[c.v]) continue;
            merged[c.u] = merged[c.v] = 1;
            map[c.u] = map[c.v] = cur++;
        }
        for (int c = 0; c < nl; ++c)
            if (!merged[c]) map[c] = cur++;
        for (long u = 0; u < N; ++u) labels[u] = map[labels[u]];
        nl = cur;
    }
    *nlabels = nl;
    return GM_OK;
}

}  // extern "C"
