// Host model of the batched solver's memory accesses (gm_batch.hip), run on
// the CPU over given image sizes and grids.  It walks every launch the way
// its kernel does -- workgroup -> range -> items -> images, tile origins, the
// LDS window ring and its carry from the tile above -- with the kernel's own
// index functions (gm_batch_plan.h), and checks:
//   * every slot a range stores and every slot a fold reads lies in its
//     buffer, and a fold reads only ranges that touch the image;
//   * every LDS ring cell a pixel's product reads holds that pixel's
//     neighbour (each stored cell is tagged with the window cell it holds);
//   * every vector index is inside its image, every 16-byte band offset fits
//     32 bits, and a system's block of vectors stays below 2^31 bytes;
//   * per launch, each pixel of each image is advanced (x, r) exactly once by
//     the fused step and is an output cell exactly once; the true residual
//     and the update-only chunks cover each pixel exactly once.
//
//   batch_index_model <nrhs> <cus> <per_cu...> -- W1 H1 W2 H2 ...
// Exit status 0 when every check holds; the first failures are printed.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../../gpumatting_amd/csrc/gm_batch_plan.h"

using namespace gm::bplan;

constexpr int kBlock = 256, kKeep = 4, kRTX = 128, kChunk = 4 * kBlock, kNQ = 5, kMaxRhs = 8, kMaxGrid = 16384;
constexpr int kSysVecs = 9;  // vectors per system block (x, r0, r1, p0, p1, Ap0, Ap1, d, b)

static long g_checks = 0, g_fail = 0;
#define REQ(c, ...)                                     \
    do {                                                \
        ++g_checks;                                     \
        if (!(c)) {                                     \
            if (g_fail++ < 20) {                        \
                fprintf(stderr, "check failed: %s: ", #c); \
                fprintf(stderr, __VA_ARGS__);           \
                fprintf(stderr, "\n");                  \
            }                                           \
        }                                               \
    } while (0)

static int batch_grid(int cus, long items, int per_cu) {
    long g = std::min((long)cus * per_cu, items);
    g = std::max(g, 1L);
    g = (g + kXcd - 1) / kXcd * kXcd;
    return (int)std::min<long>(g, kMaxGrid);
}

struct Tag {
    int gy, gx;
};

// The walk of a tiled kernel (kb_fused with TX-wide tiles, or kb_trueres with
// 128 x 2 tiles): slots of `stride` doubles per image, window ring checks,
// ownership counts.
// Threads: tile row tid / LW, column tid % LW (LW = TX + 4 for the fused
// step, whose last 4 columns' threads own no pixel; TX for the true residual).
static void walk_tiles(const std::vector<int>& W, const std::vector<int>& H, const std::vector<ImgLists>& li,
                       const Part& P, int TX, int TY, int LW, bool fused, int stride, size_t slots,
                       std::vector<std::vector<int>>& adv, std::vector<std::vector<int>>& out) {
    const int WX = TX + 4, WY = TY + 4, NEW = (WY - kKeep) * WX, KEEP = kKeep * WX;
    REQ(TY * LW == kBlock, "tile %d x %d with %d lanes", TX, TY, LW);
    std::vector<char> seen(P.G, 0);
    std::vector<Tag> ring(WY * WX);
    for (int b = 0; b < P.G; ++b) {
        const int r = range_of(b, P.G);
        REQ(r >= 0 && r < P.G && !seen[r], "workgroup %d -> range %d of %d", b, r, P.G);
        if (r < 0 || r >= P.G) continue;
        seen[r] = 1;
        const int tb = part_begin(r, P.T, P.G), te = part_begin(r + 1, P.T, P.G);
        REQ(tb >= 0 && te <= P.T && tb <= te, "range %d: [%d, %d) of %d", r, tb, te, P.T);
        int cur = -1, wx0 = -1, wy0 = 0, top = 0;
        auto flush = [&](int im) {
            const long long s0 = slot_index(r, P.K, im, P.wimg[r], stride, 0);
            REQ(im - P.wimg[r] >= 0 && im - P.wimg[r] < P.K, "range %d image %d first %d K %d", r, im, P.wimg[r], P.K);
            REQ(s0 >= 0 && (size_t)(s0 + stride) <= slots, "range %d image %d slot %lld of %zu", r, im, s0, slots);
        };
        for (int t = tb; t < te; ++t) {
            const int im = P.img[t];
            if (im != cur) {
                if (cur >= 0) flush(cur);
                cur = im;
                wx0 = -1;
            }
            const int Wi = W[im], Hi = H[im];
            const long long N = (long long)Wi * Hi;
            int x0, y0;
            tile_origin(t - (fused ? li[im].t0 : li[im].q0), fused ? li[im].nty : li[im].nqy, (Wi + TX - 1) / TX,
                        seg_len(P.T, P.G), TX, TY, x0, y0);
            REQ(x0 >= 0 && x0 < Wi && y0 >= 0 && y0 < Hi && x0 % TX == 0 && y0 % TY == 0,
                "tile %d of image %d at (%d, %d) in %dx%d", t, im, x0, y0, Wi, Hi);
            const bool carry = x0 == wx0 && y0 == wy0 + TY;
            if (carry) {
                // the rows carried over must hold the previous tile's rows TY..
                for (int wy = 0; wy < kKeep; ++wy)
                    for (int wx = 0; wx < WX; ++wx) {
                        const Tag& tg = ring[ring_row(ring_row(top + TY, WY) + wy, WY) * WX + wx];
                        REQ(tg.gy == y0 - 2 + wy && tg.gx == x0 - 2 + wx, "carried cell (%d, %d) holds (%d, %d)",
                            y0 - 2 + wy, x0 - 2 + wx, tg.gy, tg.gx);
                    }
            }
            top = carry ? ring_row(top + TY, WY) : 0;
            REQ(top >= 0 && top < WY, "ring top %d", top);
            auto store = [&](int wy, int wx) {
                const int li2 = ring_row(top + wy, WY) * WX + wx;
                REQ(li2 >= 0 && li2 < WY * WX, "LDS cell %d of %d", li2, WY * WX);
                if (li2 < 0 || li2 >= WY * WX) return;
                ring[li2] = Tag{y0 - 2 + wy, x0 - 2 + wx};
                const long long g = (long long)(y0 - 2 + wy) * Wi + (x0 - 2 + wx);
                const bool in = g >= 0 && g < N;
                if (in) REQ(4ull * (unsigned long long)g < (1ull << 32), "byte offset of %lld", g);
                // kb_fused addresses a system's vectors as byte offsets into its
                // block of kSysVecs vectors of `pitch` floats: the vector's
                // offset plus the element's must stay inside the block, which
                // must stay below 2^31 bytes (the sentinel offset 2^31 reads 0)
                if (in) {
                    const unsigned long long pitch = ((unsigned long long)N + 63) / 64 * 64;
                    REQ(4ull * kSysVecs * pitch < (1ull << 31), "system block of %llu bytes", 4ull * kSysVecs * pitch);
                    REQ(4ull * ((kSysVecs - 1) * pitch + (unsigned long long)g) < 4ull * kSysVecs * pitch,
                        "element %lld beyond its vector", g);
                }
                if (fused && advances(wy, wx, x0, y0, TX, Wi, Hi)) {
                    REQ(in, "advanced cell %lld outside image %d", g, im);
                    if (in) adv[im][g] += 1;
                }
            };
            if (!carry)
                for (int e = 0; e < KEEP; ++e) store(e / WX, e % WX);
            for (int e = 0; e < NEW; ++e) store(kKeep + e / WX, e % WX);
            wx0 = x0;
            wy0 = y0;
            for (int tid = 0; tid < kBlock; ++tid) {
                const int lane = tid % LW, row = tid / LW, x = x0 + lane, y = y0 + row;
                if (!(lane < TX && x < Wi && y < Hi)) continue;
                const long long i = (long long)y * Wi + x;
                out[im][i] += 1;
                REQ(16ull * (unsigned long long)i < (1ull << 32), "band offset of pixel %lld", i);
                for (int s = 0; s < 5; ++s)
                    for (int dx = 0; dx < 5; ++dx) {
                        const int li2 = ring_row(top + row + s, WY) * WX + lane + dx;
                        REQ(li2 >= 0 && li2 < WY * WX, "product reads LDS %d", li2);
                        if (li2 < 0 || li2 >= WY * WX) continue;
                        const Tag& tg = ring[li2];
                        REQ(tg.gy == y - 2 + s && tg.gx == x - 2 + dx, "pixel (%d, %d) reads (%d, %d) for (%d, %d)", y,
                            x, tg.gy, tg.gx, y - 2 + s, x - 2 + dx);
                    }
            }
        }
        if (cur >= 0) flush(cur);
    }
}

// A fold of image im's slots over ranges [wf, wl] (img_fold).
static void check_fold(const Part& P, int im, int stride, int off, int nq, size_t slots) {
    const int wf = P.wf[im], wl = P.wl[im];
    REQ(wf <= wl, "image %d is covered by no range", im);
    for (int r = wf; r <= wl; ++r) {
        REQ(im - P.wimg[r] >= 0 && im - P.wimg[r] < P.K, "fold of image %d reads range %d (first image %d, K %d)", im,
            r, P.wimg[r], P.K);
        const long long s = slot_index(r, P.K, im, P.wimg[r], stride, off);
        REQ(s >= 0 && (size_t)(s + nq) <= slots, "fold of image %d range %d slot %lld of %zu", im, r, s, slots);
        // the range does touch image im, or is empty (its slot then stays 0)
        const int tb = part_begin(r, P.T, P.G), te = part_begin(r + 1, P.T, P.G);
        bool touch = tb >= te;
        for (int t = tb; t < te && !touch; ++t) touch = P.img[t] == im;
        REQ(touch, "range %d folded for image %d does not touch it", r, im);
    }
}

// One model run; fused tiles of window width WX (TX = WX - 4 pixels, kBlock / WX rows).
static void model(const std::vector<int>& W, const std::vector<int>& H, int nrhs, int cus, int per_cu, int WX) {
    const int nimg = (int)W.size(), NR = nrhs <= 1 ? 1 : nrhs <= 2 ? 2 : nrhs <= 4 ? 4 : 8;
    const int TX = WX - 4, TY = kBlock / WX;
    std::vector<ImgLists> li;
    Part pf, pq, pc;
    build_lists(W, H, TX, TY, kRTX, kBlock / kRTX, kChunk, li, pf, pq, pc);
    split(pf, batch_grid(cus, (long)pf.img.size(), per_cu), nimg);
    split(pq, batch_grid(cus, (long)pq.img.size(), 16), nimg);
    split(pc, batch_grid(cus, (long)pc.img.size(), 32), nimg);
    const size_t fhalf = (size_t)pf.G * pf.K * kNQ * NR;
    const size_t s3 = (size_t)pq.G * pq.K * 4 * kMaxRhs;
    std::vector<std::vector<int>> adv(nimg), out(nimg), out3(nimg), adv3(nimg), upd(nimg);
    for (int i = 0; i < nimg; ++i) {
        const size_t N = (size_t)W[i] * H[i];
        adv[i].assign(N, 0);
        out[i].assign(N, 0);
        adv3[i].assign(N, 0);
        out3[i].assign(N, 0);
        upd[i].assign(N, 0);
    }
    // kb_fused (one launch)
    walk_tiles(W, H, li, pf, TX, TY, WX, true, kNQ * NR, fhalf, adv, out);
    // kb_trueres (128 x 2 tiles, 4 sums per system of kMaxRhs)
    walk_tiles(W, H, li, pq, kRTX, kBlock / kRTX, kRTX, false, 4 * kMaxRhs, s3, adv3, out3);
    // kb_fupdate: chunks of kChunk pixels, 4 per thread
    for (int b = 0; b < pc.G; ++b) {
        const int r = range_of(b, pc.G);
        const int cb = part_begin(r, pc.T, pc.G), ce = part_begin(r + 1, pc.T, pc.G);
        for (int c = cb; c < ce; ++c) {
            const int im = pc.img[c];
            const long long N = (long long)W[im] * H[im];
            for (int tid = 0; tid < kBlock; ++tid) {
                const long long i0 = (long long)(c - li[im].c0) * kChunk + 4 * tid;
                if (i0 >= N) continue;
                for (long long i = i0; i < std::min(i0 + 4, N); ++i) upd[im][i] += 1;
            }
        }
    }
    for (int im = 0; im < nimg; ++im) {
        for (int j = 0; j < NR; ++j) check_fold(pf, im, kNQ * NR, kNQ * j, kNQ, fhalf);  // kb_fused, kb_fupdate, kb_ffold
        for (int j = 0; j < kMaxRhs; ++j) check_fold(pq, im, 4 * kMaxRhs, 4 * j, 4, s3);  // kb_trueres_fin
        REQ(pc.wf[im] <= pc.wl[im], "image %d has no update chunk range", im);
        for (size_t i = 0; i < adv[im].size(); ++i) {
            REQ(adv[im][i] == 1, "image %d (%dx%d) pixel %zu advanced %d times (TX %d, G %d)", im, W[im], H[im], i,
                adv[im][i], TX, pf.G);
            REQ(out[im][i] == 1, "image %d pixel %zu is an output cell %d times", im, i, out[im][i]);
            REQ(out3[im][i] == 1, "image %d pixel %zu true residual %d times", im, i, out3[im][i]);
            REQ(upd[im][i] == 1, "image %d pixel %zu updated %d times by the chunks", im, i, upd[im][i]);
        }
    }
}

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s nrhs cus per_cu... -- W1 H1 ...\n", argv[0]);
        return 2;
    }
    const int nrhs = atoi(argv[1]), cus = atoi(argv[2]);
    std::vector<int> pcs, W, H;
    int a = 3;
    for (; a < argc && strcmp(argv[a], "--"); ++a) pcs.push_back(atoi(argv[a]));
    for (++a; a + 1 < argc; a += 2) {
        W.push_back(atoi(argv[a]));
        H.push_back(atoi(argv[a + 1]));
    }
    if (W.empty() || pcs.empty()) return 2;
    for (int WX : {64, 128, 256})
        for (int pc : pcs) model(W, H, nrhs, cus, pc, WX);
    printf("%ld checks, %ld failed (%zu images, nrhs %d, %zu grids x 3 tile widths)\n", g_checks, g_fail, W.size(),
           nrhs, pcs.size());
    return g_fail ? 1 : 0;
}
