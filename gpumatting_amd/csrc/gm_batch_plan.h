// gm_batch_plan.h -- work lists, partitions and index arithmetic of the
// batched solver (gm_batch.hip), in plain C++ so that the kernels and a host
// model of their memory accesses share one definition.
//
// A work list concatenates the items (tiles or chunks) of all images; a
// launch splits it over G ranges, one per workgroup, contiguous and in list
// order.  Workgroup b takes range range_of(b, G): the XCD of b (b % 8) takes
// blocks of kXcdBlock consecutive ranges, dealt to the XCDs in turn, so
// neighbouring ranges -- which share window rows at their ends -- mostly run
// on one XCD and meet in its L2, while an image that spans many ranges is
// spread over every XCD (late in a solve only a few images still step).  A range
// that touches several images keeps one slot of partial sums per image:
// slot (r * K + im - wimg[r]), K = the most images one range touches.
#pragma once

#include <stddef.h>

#include <algorithm>
#include <vector>

#ifdef __HIPCC__
#define GM_HD __host__ __device__
#else
#define GM_HD
#endif

namespace gm {
namespace bplan {

constexpr int kXcd = 8;

constexpr int kXcdBlock = 8;
// A bijection of [0, G) when G % kXcd == 0 (else the identity): XCD x's l-th
// workgroup takes range (l / B * kXcd + x) * B + l % B of the first
// kXcd * B * (M / B) ranges (M = G / kXcd, B = kXcdBlock), and the XCDs
// share the M % B ranges left at the end in contiguous runs.
GM_HD inline int range_of(int b, int G) {
    if (G % kXcd) return b;
    const int M = G / kXcd, x = b % kXcd, l = b / kXcd, full = M / kXcdBlock * kXcdBlock;
    if (l < full) return (l / kXcdBlock * kXcd + x) * kXcdBlock + l % kXcdBlock;
    const int rem = M - full;
    return kXcd * full + x * rem + (l - full);
}
GM_HD inline int part_begin(int r, int T, int G) { return (int)(((long long)r * T) / G); }
GM_HD inline long long slot_index(int r, int K, int im, int wimg_r, int stride, int off) {
    return ((long long)r * K + (im - wimg_r)) * stride + off;
}
// Origin of local tile lt of an image of ncol TX-wide columns of nty tiles.
// The tiles are numbered in segments of `seg` tile rows: segment s holds tile
// rows [s*seg, min((s+1)*seg, nty)) of every column, column after column, so
// that consecutive ranges (one XCD, resident together) walk side by side down
// neighbouring columns and the cache lines their windows share meet in L2.
GM_HD inline void tile_origin(int lt, int nty, int ncol, int seg, int TX, int TY, int& x0, int& y0) {
    const int per = seg * ncol;  // tiles of a full segment
    const int s = lt / per, rem = lt - s * per;
    const int len = nty - s * seg < seg ? nty - s * seg : seg;
    const int c = rem / len;
    x0 = c * TX;
    y0 = (s * seg + rem - c * len) * TY;
}
// Segment length for a list of T tiles over G ranges: a range's worth of
// tiles (at least 2 tile rows).
GM_HD inline int seg_len(int T, int G) { return T / G > 2 ? T / G : 2; }
GM_HD inline int ring_row(int s, int WY) { return s >= WY ? s - WY : s; }

// Whether window cell (wy, wx) of the fused step's tile at (x0, y0) (window
// rows 0..TY+3 are image rows y0-2.., columns 0..TX+3 are x0-2..) is one the
// tile advances (x and r): a pixel of the image in the tile's own columns, in
// a new window row (wy >= 4), or in rows 0 and 1 on the image's first tile
// row (the only tiles that see them).  Going down a column every pixel is
// then advanced exactly once.
GM_HD inline bool advances(int wy, int wx, int x0, int y0, int TX, int W, int H) {
    const int gy = y0 - 2 + wy, gx = x0 - 2 + wx;
    if (gy < 0 || gy >= H || wx < 2 || wx >= TX + 2 || gx >= W) return false;
    return wy >= 4 || y0 == 0;
}

// Host side of one partition: items -> image, ranges -> first image, and per
// image the first / last range covering it.
struct Part {
    int T = 0, G = 0, K = 1;
    std::vector<int> img;   // item -> image (image-sorted)
    std::vector<int> wimg;  // range -> image of its first item (of item min(begin, T - 1) if empty)
    std::vector<int> wf, wl;  // image -> first / last non-empty range covering it (wf > wl: none)
};

// Split p.img (filled) over G ranges.
inline void split(Part& p, int G, int nimg) {
    p.T = (int)p.img.size();
    p.G = G;
    p.K = 1;
    p.wimg.assign(G, 0);
    p.wf.assign(nimg, G);
    p.wl.assign(nimg, -1);
    for (int r = 0; r < G; ++r) {
        const int tb = part_begin(r, p.T, G), te = part_begin(r + 1, p.T, G);
        if (tb >= te) {
            p.wimg[r] = p.T ? p.img[std::min(tb, p.T - 1)] : 0;
            continue;
        }
        const int i0 = p.img[tb], i1 = p.img[te - 1];
        p.wimg[r] = i0;
        p.K = std::max(p.K, i1 - i0 + 1);
        for (int im = i0; im <= i1; ++im) {
            p.wf[im] = std::min(p.wf[im], r);
            p.wl[im] = std::max(p.wl[im], r);
        }
    }
}

// Per image: where its items start in the three work lists.
struct ImgLists {
    int t0, nty;  // fused tiles, numbered down TX-wide columns; tiles per column
    int q0, nqy;  // true-residual tiles
    int c0;       // update-only chunks of `chunk` pixels
};

// The fused-tile (TX x TY), true-residual-tile (RTX x RTY) and chunk lists of
// images W[i] x H[i].
inline void build_lists(const std::vector<int>& W, const std::vector<int>& H, int TX, int TY, int RTX, int RTY,
                        int chunk, std::vector<ImgLists>& li, Part& pf, Part& pq, Part& pc) {
    const int n = (int)W.size();
    li.assign(n, ImgLists{});
    pf.img.clear();
    pq.img.clear();
    pc.img.clear();
    for (int i = 0; i < n; ++i) {
        ImgLists& L = li[i];
        const long long N = (long long)W[i] * H[i];
        L.nty = (H[i] + TY - 1) / TY;
        L.t0 = (int)pf.img.size();
        pf.img.insert(pf.img.end(), (size_t)L.nty * ((W[i] + TX - 1) / TX), i);
        L.nqy = (H[i] + RTY - 1) / RTY;
        L.q0 = (int)pq.img.size();
        pq.img.insert(pq.img.end(), (size_t)L.nqy * ((W[i] + RTX - 1) / RTX), i);
        L.c0 = (int)pc.img.size();
        pc.img.insert(pc.img.end(), (size_t)((N + chunk - 1) / chunk), i);
    }
}

}  // namespace bplan
}  // namespace gm
