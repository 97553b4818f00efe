// gm_batch.hip -- many images of uneven size, several right-hand sides each,
// solved together (no reference counterpart).
//
// Each image i has one Levin Laplacian L_i (25 FP32 bands, no data term) and
// nrhs independent constraint sets j, i.e. nrhs systems
//     (L_i + diag(lam_ij)) x_ij = lam_ij * t_ij .
// The systems of one image share L_i's off-diagonal bands: one pass over a
// tile loads those 24 bands into registers once and applies them to every
// right-hand side; each system brings its own FP32 diagonal d_ij = L_ii +
// lam_ij.  Jacobi-PCG per system, with the single-system kernels' sign,
// residual-replacement and hand-off conventions (see Scalars).
//
// Work lists.  The step and true-residual kernels walk 128x2-pixel tiles (one
// pixel per thread, any image width), numbered down each image's 128-pixel
// columns; the update kernel walks 1024-pixel chunks.  Each list concatenates
// all images.  Going down a column, a tile's x window (2-pixel halo) shares 4
// of its 6 rows with the tile above: the window is an LDS ring, and a
// workgroup that has just done the tile above loads only the 2 new rows.
// (Against 64x4 tiles, the 128-wide window rows touch fewer cache lines
// outside the tile per pixel.)  A launch splits its list into
// contiguous ranges, one per workgroup, so a workgroup touches few images: it
// accumulates its partial sums per (image, system) in registers and stores
// one slot per image it touches.  The consumer kernel folds an image's slots
// -- those of the workgroups whose ranges cover the image, in index order --
// with no fence or atomic (the kernel boundary orders them).  Systems that
// have converged are skipped.  Iterations are captured in a hipGraph (kernel
// nodes only) and replayed.
//
// Mixed precision: the off-diagonal bands are also kept in FP16 or BF16
// (unscaled; |L_ij| < 20), the per-system diagonals stay FP32, and the residual is
// replaced by the FP64-accumulated true residual (FP32 L) every 50
// iterations (see kDrift2 for when the search direction is kept).
#include <string.h>

#include <algorithm>
#include <vector>

#include "gm_internal.h"

namespace gm {

gm_status levin_assemble_raw(float* A, size_t apitch, float* b, float* dinv, const float* im, int nch, size_t pitch,
                             const float* lam, const float* tgt, const unsigned char* scr, float gamma, float eps,
                             int mode, int W, int H, gm_stream_t s);

constexpr int kMaxRhs = 8;
constexpr int kTX = 128, kTY = kBlock / kTX;  // step / true-residual tile: 128 x 2 pixels
constexpr int kWX = kTX + 4, kWY = kTY + 4;   // staged window (2-pixel halo), kWY rows
constexpr int kKeep = 4;                      // window rows a tile shares with the tile directly above
constexpr int kNew = (kWY - kKeep) * kWX;     // cells of the window rows below those
constexpr int kNewPer = (kNew + kBlock - 1) / kBlock;
static_assert(kTY * kTX == kBlock && kTY <= kKeep, "one pixel per thread; a tile shares 4 window rows");
constexpr int kChunk = 4 * kBlock;            // update-kernel chunk: 1024 pixels, 4 per thread
// The window is a ring of kWY LDS rows: window row wy of a tile is ring row
// ring_row(top + wy); the tile below starts at ring_row(top + kTY), so the
// kKeep rows it shares stay where they are.  (s < 2 kWY)
__device__ __forceinline__ int ring_row(int s) { return s >= kWY ? s - kWY : s; }

// Static split of a work list over the workgroups of a launch.
struct BPart {
    int T;            // items in the list
    int G;            // workgroups
    int K;            // slots per workgroup: the most images one range touches
    const int* img;   // item -> image
    const int* wimg;  // workgroup -> image of its first item
};
__device__ __forceinline__ int part_begin(const BPart& P, int b) { return (int)(((long long)b * P.T) / P.G); }

struct BImg {
    int W, H, N, sys0, nsys;
    int t0, nt, nty;  // step tiles: first index in the tile list, count, tiles per image column
    int c0, nc;       // update chunks
    int wf1, wl1;     // first / last workgroup of the tile partition covering the image
    int wf2, wl2;     // same for the chunk partition
    const float* L;
    size_t pitch;
    const unsigned short* L16;
    size_t pitch16;
    const float* Lg;  // FP32 form: the off-diagonal bands 4 per pixel in 6 groups (kb_group32)
};

struct BSys {
    float *x, *r, *p0, *p1, *Ap, *d, *dinv, *b;
    float* z;  // D^-1 r, kept beside r by every kernel that writes r
};

// Sum an image's slots: slot[(b * K + im - wimg[b]) * stride + off] over the
// workgroups b in [wf, wl], in a fixed order; every thread gets the total.
__device__ __forceinline__ double img_fold(const double* __restrict__ slot, const BPart& P, int im, int wf, int wl,
                                           int stride, int off) {
    __shared__ double fr[kWaves];
    __shared__ double ft;
    double s = 0;
    for (int b = wf + (int)threadIdx.x; b <= wl; b += kBlock)
        s += slot[((size_t)b * P.K + (im - P.wimg[b])) * stride + off];
    s = wave_sum(s);
    __syncthreads();  // an earlier call's readers are done with fr / ft
    if ((threadIdx.x & 63) == 0) fr[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += fr[w];
        ft = t;
    }
    __syncthreads();
    return ft;
}

// Store this workgroup's per-system partial sums for one image (block
// reduction of acc[j][q], j < kMaxRhs, q < NQ) and clear them.
template <int NQ>
__device__ __forceinline__ void flush_slot(double (&acc)[kMaxRhs][NQ], double* __restrict__ dst) {
    __shared__ double red[kMaxRhs * NQ][kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kMaxRhs; ++j)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double s = wave_sum(acc[j][q]);
            if (lane == 0) red[j * NQ + q][wid] = s;
            acc[j][q] = 0.0;
        }
    __syncthreads();
    if (threadIdx.x < kMaxRhs * NQ) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += red[threadIdx.x][w];
        dst[threadIdx.x] = t;
    }
    __syncthreads();  // red is reused by the next flush
}

// Step-kernel view of one image's systems: fold pending update partials and
// run the convergence test (every thread; results to chand).
__device__ __forceinline__ void b_step_enter(const BImg& I, int im, const Scalars* sc,
                                             const double* __restrict__ slot2, const BPart& P2, int verify,
                                             Hand* chand) {
    __syncthreads();  // chand of the previous image is no longer read
    for (int j = 0; j < I.nsys; ++j) {
        const Scalars* s = &sc[I.sys0 + j];
        Hand h{s->rz, s->rz_prev, s->rr, s->beta, s->iters, kHandRun, kRun, 0, 0};
        if (s->done != kRun || s->verify) {
            h.mode = kHandIdle;
        } else if (s->pend) {
            const double rz = img_fold(slot2, P2, im, I.wf2, I.wl2, 2 * kMaxRhs, 2 * j);
            const double rr = img_fold(slot2, P2, im, I.wf2, I.wl2, 2 * kMaxRhs, 2 * j + 1);
            if (!isfinite(rz) || !isfinite(rr)) {
                h.mode = kHandStop;
                h.done = kNonFinite;
            } else {
                h.rz_prev = h.rz;
                h.beta = (float)(rz / h.rz);
                h.rz = rz;
                h.rr = rr;
                if (s->tol2 >= 0 && rr <= s->tol2) {
                    h.mode = kHandStop;
                    if (verify) h.verify = 1;
                    else h.done = kConverged;
                } else if (h.iters >= s->maxit) {
                    h.mode = kHandStop;
                    h.done = kMaxIt;
                }
            }
        }
        if (threadIdx.x == 0) chand[j] = h;
    }
    __syncthreads();
}

// Element at a byte offset from a vector's base, in global memory (the
// pointers come from device tables, so the compiler cannot tell by itself and
// would use flat accesses, which also count against the LDS wait counter).
// Images have fewer than 2^28 pixels, so byte offsets of up to 16 bytes per
// pixel fit 32 bits: the access takes
// the base from scalar registers and one 32-bit offset register, which the
// systems' vectors share.
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ float at_off(const float* base, unsigned off) {
    return *reinterpret_cast<const __attribute__((address_space(1))) float*>((const gchar*)base + off);
}
__device__ __forceinline__ void store_off(float* base, unsigned off, float v) {
    *reinterpret_cast<__attribute__((address_space(1))) float*>((gchar*)base + off) = v;
}
__device__ __forceinline__ f32x4 at_off4(const float* base, unsigned off) {
    return *reinterpret_cast<const __attribute__((address_space(1))) f32x4*>((const gchar*)base + off);
}
__device__ __forceinline__ void store_off4(float* base, unsigned off, f32x4 v) {
    *reinterpret_cast<__attribute__((address_space(1))) f32x4*>((gchar*)base + off) = v;
}
template <class T>
__device__ __forceinline__ T nt_at_off(const T* base, unsigned off) {
    return __builtin_nontemporal_load(reinterpret_cast<const __attribute__((address_space(1))) T*>((const gchar*)base + off));
}

// The 24 off-diagonal band values of pixel i (band order, 12 skipped), as
// loaded: FP32 values, or the 16-bit bands in groups of 8 per pixel (kb_to16).
template <int FMT>
struct PixBands {
    float a[24];
};
template <>
struct PixBands<kF16> {
    u32x4 g[3];
};
template <>
struct PixBands<kBF16> {
    u32x4 g[3];
};
// the FP32 bands in groups of 4 per pixel (kb_group32): six 16-byte loads
constexpr int kF32G = 3;
template <>
struct PixBands<kF32G> {
    f32x4 g[6];
};
template <int FMT>
__device__ __forceinline__ void load_pixel_bands(const BImg& I, int i, PixBands<FMT>& pb) {
    if constexpr (FMT == kF32) {
#pragma unroll
        for (int k = 0, kk = 0; k < 25; ++k) {
            if (k == 12) continue;
            pb.a[kk++] = nt_at_off(I.L + (size_t)k * I.pitch, 4u * (unsigned)i);
        }
    } else if constexpr (FMT == kF32G) {
        const f32x4* g = reinterpret_cast<const f32x4*>(I.Lg);
#pragma unroll
        for (int q = 0; q < 6; ++q) pb.g[q] = nt_at_off(g + (size_t)q * I.pitch, 16u * (unsigned)i);
    } else {
        const u32x4* g = reinterpret_cast<const u32x4*>(I.L16);
#pragma unroll
        for (int q = 0; q < 3; ++q) pb.g[q] = nt_at_off(g + (size_t)q * I.pitch16, 16u * (unsigned)i);
    }
}
template <int FMT>
__device__ __forceinline__ void pixel_bands(const PixBands<FMT>& pb, float (&a)[24]) {
    if constexpr (FMT == kF32) {
#pragma unroll
        for (int k = 0; k < 24; ++k) a[k] = pb.a[k];
    } else if constexpr (FMT == kF32G) {
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int h = 0; h < 4; ++h) a[4 * q + h] = pb.g[q][h];
    } else {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const unsigned short u = (unsigned short)(pb.g[q][h >> 1] >> (16 * (h & 1)));
                a[8 * q + h] = FMT == kBF16 ? bf16_to_f32(u) : (float)__builtin_bit_cast(_Float16, u);
            }
    }
}

// Window staging of the step and true-residual kernels.  cell(j, o) is
// system j's staged value at the matrix row of byte offset o; cells beside
// the image alias the neighbouring row, exactly like the row-wise product
// (their band entries are 0), and cells outside the vector are 0.
// window_new_load: the cells of window rows kKeep.. into registers (issued
// before the previous tile is done with the LDS ring).
template <class Cell>
__device__ __forceinline__ void window_new_load(const BImg& I, unsigned mask, int x0, int y0, Cell cell,
                                                float (&sv)[kMaxRhs][kNewPer]) {
#pragma unroll
    for (int j = 0; j < kMaxRhs; ++j) {
        if (!((mask >> j) & 1u)) continue;
#pragma unroll
        for (int q = 0; q < kNewPer; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            const int wy = kKeep + e / kWX, wx = e % kWX;
            const long long g = (long long)(y0 - 2 + wy) * I.W + (x0 - 2 + wx);
            sv[j][q] = (e < kNew && g >= 0 && g < I.N) ? cell(j, 4u * (unsigned)g) : 0.f;
        }
    }
}
// window_store: the ring rows of the tile at (x0, y0) with its first row at
// ring row top: the registers of window_new_load, and, for a tile that is not
// directly below the previous one (carry false), its first kKeep rows loaded
// here.  Every thread calls it (two barriers).
template <class Cell>
__device__ __forceinline__ void window_store(const BImg& I, unsigned mask, int x0, int y0, bool carry, int top,
                                             Cell cell, const float (&sv)[kMaxRhs][kNewPer],
                                             float (*win)[kWY * kWX]) {
    __syncthreads();  // the previous tile is done with win
    if (!carry) {
        for (int j = 0; j < kMaxRhs; ++j) {
            if (!((mask >> j) & 1u)) continue;
            for (int e = (int)threadIdx.x; e < kKeep * kWX; e += kBlock) {
                const int wy = e / kWX, wx = e - wy * kWX;
                const long long g = (long long)(y0 - 2 + wy) * I.W + (x0 - 2 + wx);
                win[j][ring_row(top + wy) * kWX + wx] = (g >= 0 && g < I.N) ? cell(j, 4u * (unsigned)g) : 0.f;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kMaxRhs; ++j) {
        if (!((mask >> j) & 1u)) continue;
#pragma unroll
        for (int q = 0; q < kNewPer; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            if (e < kNew) {
                const int r = e / kWX;
                win[j][ring_row(top + kKeep + r) * kWX + (e - r * kWX)] = sv[j][q];
            }
        }
    }
    __syncthreads();
}

// K1: per running system p = -D^-1 r + beta p_old (staged), Ap = (L_off +
// diag d) p, <p, Ap> partials.
// (at most 128 VGPRs: 4 waves per SIMD)
template <int FMT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void kb_step(const BImg* __restrict__ imgs, const BSys* __restrict__ sys,
                                                  Scalars* sc, BPart P1, BPart P2, double* __restrict__ slot1,
                                                  const double* __restrict__ slot2, int verify) {
    __shared__ float win[kMaxRhs][kWY * kWX];
    __shared__ Hand chand[kMaxRhs];
    const int b = blockIdx.x;
    const int tb = part_begin(P1, b), te = part_begin(P1, b + 1);
    const int lane = (int)threadIdx.x % kTX, row = (int)threadIdx.x / kTX;
    double acc[kMaxRhs][1] = {};
    int cur = -1;
    unsigned mask = 0;
    BImg I{};
    // the window in LDS: origin of its tile (wx0 < 0: none) and ring row of window row 0
    int wx0 = -1, wy0 = 0, top = 0;
    auto cell = [&](int j, unsigned o) {
        const BSys& S = sys[I.sys0 + j];
        const float beta = chand[j].beta;
        const float* __restrict__ pold = (chand[j].iters & 1) ? S.p1 : S.p0;
        const float z = at_off(S.z, o);
        return beta != 0.f ? fmaf(beta, at_off(pold, o), -z) : -z;
    };
    for (int t = tb; t < te; ++t) {
        const int im = P1.img[t];
        if (im != cur) {
            if (cur >= 0) flush_slot<1>(acc, slot1 + ((size_t)b * P1.K + (cur - P1.wimg[b])) * kMaxRhs);
            I = imgs[im];
            b_step_enter(I, im, sc, slot2, P2, verify, chand);
            if (b == I.wf1 && threadIdx.x == 0)
                for (int j = 0; j < I.nsys; ++j) sc[I.sys0 + j].hand = chand[j];
            mask = 0;
            for (int j = 0; j < I.nsys; ++j)
                if (chand[j].mode == kHandRun) mask |= 1u << j;
            cur = im;
            wx0 = -1;
        }
        if (mask == 0) continue;
        const int lt = t - I.t0;
        const int x0 = (lt / I.nty) * kTX, y0 = (lt % I.nty) * kTY;
        const int x = x0 + lane, y = y0 + row;
        const bool valid = x < I.W && y < I.H;
        const int i = y * I.W + x;
        // the tile directly above is in LDS: its window rows kTY.. are this one's rows 0..
        const bool carry = x0 == wx0 && y0 == wy0 + kTY;
        top = carry ? ring_row(top + kTY) : 0;
        // every global load of the tile first, then LDS
        PixBands<FMT> pb;
        if (valid) load_pixel_bands<FMT>(I, i, pb);
        float dd[kMaxRhs];  // each system's diagonal at the thread's pixel
#pragma unroll
        for (int j = 0; j < kMaxRhs; ++j)
            if ((mask >> j) & 1u) dd[j] = valid ? at_off(sys[I.sys0 + j].d, 4u * (unsigned)i) : 0.f;
        float sv[kMaxRhs][kNewPer];
        window_new_load(I, mask, x0, y0, cell, sv);
        window_store(I, mask, x0, y0, carry, top, cell, sv, win);
        wx0 = x0;
        wy0 = y0;
        if (!valid) continue;
        float a[24];
        pixel_bands<FMT>(pb, a);
        // LDS offsets of the thread's 5 window rows (cell (row + q, lane) is pixel (y - 2 + q, x - 2))
        int rb[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) rb[q] = ring_row(top + row + q) * kWX + lane;
#pragma unroll
        for (int j = 0; j < kMaxRhs; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const BSys S = sys[I.sys0 + j];
            const float* __restrict__ w = win[j];
            float yv = 0.f;
#pragma unroll
            for (int k = 0, kk = 0; k < 25; ++k) {
                if (k == 12) continue;
                yv = fmaf(a[kk], w[rb[k / 5] + (k % 5)], yv);
                ++kk;
            }
            const float pc = w[rb[2] + 2];
            yv = fmaf(dd[j], pc, yv);
            float* __restrict__ pnew = (chand[j].iters & 1) ? S.p0 : S.p1;
            store_off(pnew, 4u * (unsigned)i, pc);
            store_off(S.Ap, 4u * (unsigned)i, yv);
            acc[j][0] += (double)(pc * yv);
        }
    }
    if (cur >= 0) flush_slot<1>(acc, slot1 + ((size_t)b * P1.K + (cur - P1.wimg[b])) * kMaxRhs);
}

enum { kBRun = 0, kBIdle = 10, kBStop = 11 };  // update-kernel outcome per system (else kBreakdown / kNonFinite)

// K2: per running system x += alpha p, r += alpha Ap; <r, D^-1 r>, ||r||^2 partials.
__global__ __launch_bounds__(kBlock) void kb_update(const BImg* __restrict__ imgs, const BSys* __restrict__ sys,
                                                    Scalars* sc, BPart P1, BPart P2, const double* __restrict__ slot1,
                                                    double* __restrict__ slot2) {
    __shared__ Hand chand[kMaxRhs];
    __shared__ int cst[kMaxRhs];
    __shared__ float calpha[kMaxRhs];
    __shared__ double cpap[kMaxRhs];
    const int b = blockIdx.x;
    const int cb = part_begin(P2, b), ce = part_begin(P2, b + 1);
    double acc[kMaxRhs][2] = {};
    int cur = -1;
    unsigned mask = 0;
    BImg I{};
    for (int c = cb; c < ce; ++c) {
        const int im = P2.img[c];
        if (im != cur) {
            if (cur >= 0) flush_slot<2>(acc, slot2 + ((size_t)b * P2.K + (cur - P2.wimg[b])) * 2 * kMaxRhs);
            I = imgs[im];
            __syncthreads();  // chand / cst of the previous image are no longer read
            for (int j = 0; j < I.nsys; ++j) {
                const Hand h = sc[I.sys0 + j].hand;
                int st = kBRun;
                float al = 0.f;
                double pAp = 0;
                if (h.mode == kHandIdle) {
                    st = kBIdle;
                } else if (h.mode == kHandStop) {
                    st = kBStop;
                } else {
                    pAp = img_fold(slot1, P1, im, I.wf1, I.wl1, kMaxRhs, j);
                    if (!isfinite(pAp)) st = kNonFinite;
                    else if (!(pAp > 0)) st = kBreakdown;
                    else al = (float)(h.rz / pAp);
                }
                if (threadIdx.x == 0) {
                    chand[j] = h;
                    cst[j] = st;
                    calpha[j] = al;
                    cpap[j] = pAp;
                }
            }
            __syncthreads();
            if (b == I.wf2 && threadIdx.x == 0)
                for (int j = 0; j < I.nsys; ++j) {
                    Scalars* s = &sc[I.sys0 + j];
                    const Hand& h = chand[j];
                    if (cst[j] == kBStop) {
                        publish(s, h, h.done, h.verify, h.iters, 0, 0);
                    } else if (cst[j] == kBRun) {
                        publish(s, h, kRun, 0, h.iters + 1, 1, 0);
                        s->pAp = cpap[j];
                        s->alpha = calpha[j];
                    } else if (cst[j] != kBIdle) {
                        publish(s, h, cst[j], 0, h.iters, 0, 0);
                        s->pAp = cpap[j];
                    }
                }
            mask = 0;
            for (int j = 0; j < I.nsys; ++j)
                if (cst[j] == kBRun) mask |= 1u << j;
            cur = im;
        }
        if (mask == 0) continue;
        const int i0 = (c - I.c0) * kChunk + 4 * (int)threadIdx.x;
        if (i0 >= I.N) continue;
#pragma unroll
        for (int j = 0; j < kMaxRhs; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const BSys S = sys[I.sys0 + j];
            const float al = calpha[j];
            const float* __restrict__ p = (chand[j].iters & 1) ? S.p0 : S.p1;  // the p_new K1 wrote
            if (i0 + 4 <= I.N) {  // per-system vectors are 256-byte aligned and i0 % 4 == 0
                const unsigned o = 4u * (unsigned)i0;
                const f32x4 xv = at_off4(S.x, o), rv = at_off4(S.r, o), pv = at_off4(p, o), av = at_off4(S.Ap, o),
                            dv = at_off4(S.dinv, o);
                float xs[4] = {xv.x, xv.y, xv.z, xv.w}, rs[4] = {rv.x, rv.y, rv.z, rv.w};
                const float ps[4] = {pv.x, pv.y, pv.z, pv.w}, as[4] = {av.x, av.y, av.z, av.w};
                const float ds[4] = {dv.x, dv.y, dv.z, dv.w};
                float zs[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    xs[q] = fmaf(al, ps[q], xs[q]);
                    rs[q] = fmaf(al, as[q], rs[q]);
                    zs[q] = ds[q] * rs[q];
                    acc[j][0] += (double)(rs[q] * zs[q]);
                    acc[j][1] += (double)(rs[q] * rs[q]);
                }
                store_off4(S.x, o, f32x4{xs[0], xs[1], xs[2], xs[3]});
                store_off4(S.r, o, f32x4{rs[0], rs[1], rs[2], rs[3]});
                store_off4(S.z, o, f32x4{zs[0], zs[1], zs[2], zs[3]});
            } else {
                for (int i = i0; i < I.N; ++i) {
                    const unsigned o = 4u * (unsigned)i;
                    store_off(S.x, o, fmaf(al, at_off(p, o), at_off(S.x, o)));
                    const float rn = fmaf(al, at_off(S.Ap, o), at_off(S.r, o));
                    const float zn = at_off(S.dinv, o) * rn;
                    store_off(S.r, o, rn);
                    store_off(S.z, o, zn);
                    acc[j][0] += (double)(rn * zn);
                    acc[j][1] += (double)(rn * rn);
                }
            }
        }
    }
    if (cur >= 0) flush_slot<2>(acc, slot2 + ((size_t)b * P2.K + (cur - P2.wimg[b])) * 2 * kMaxRhs);
}

// Fold pending update partials into the Scalars (one workgroup per image;
// launched before the host reads the final state).
__global__ __launch_bounds__(kBlock) void kb_fold(const BImg* __restrict__ imgs, Scalars* sc, BPart P2,
                                                  const double* __restrict__ slot2, int verify) {
    const int im = blockIdx.x;
    const BImg I = imgs[im];
    for (int j = 0; j < I.nsys; ++j) {
        Scalars* s = &sc[I.sys0 + j];
        if (s->done != kRun || s->verify || !s->pend) continue;
        const double rz = img_fold(slot2, P2, im, I.wf2, I.wl2, 2 * kMaxRhs, 2 * j);
        const double rr = img_fold(slot2, P2, im, I.wf2, I.wl2, 2 * kMaxRhs, 2 * j + 1);
        if (threadIdx.x != 0) continue;
        s->pend = 0;
        if (!isfinite(rz) || !isfinite(rr)) {
            s->done = kNonFinite;
            continue;
        }
        s->rz_prev = s->rz;
        s->beta = (float)(rz / s->rz);
        s->rz = rz;
        s->rr = rr;
        if (s->tol2 >= 0 && rr <= s->tol2) {
            if (verify) s->verify = 1;
            else s->done = kConverged;
        } else if (s->iters >= s->maxit) {
            s->done = kMaxIt;
        }
    }
}

// systems a true-residual pass selects (op: see kb_trueres)
__device__ __forceinline__ bool tr_selected(const Scalars* s, int op) {
    return op == 0 || op == 2 || (op == 1 && s->done == kRun) || (op == 3 && s->done == kRun && s->verify);
}

// FP64-accumulated true residual t = (L_off + diag d) x - b of the selected
// systems, partials (<t, D^-1 t>, ||t||^2, drift ||t - r||^2, ||b||^2) per
// image slot; kb_trueres_fin folds them.  op 0: measure only (g0); 1: replace
// r for running systems; 2: start (all systems, beta = 0, also ||b||^2); 3:
// replace r for systems whose recursive residual asked for verification.
__global__ __launch_bounds__(kBlock) void kb_trueres(const BImg* __restrict__ imgs, const BSys* __restrict__ sys,
                                                     const Scalars* sc, BPart P1, int op, double* __restrict__ slot3) {
    __shared__ float win[kMaxRhs][kWY * kWX];
    const int b = blockIdx.x;
    const int tb = part_begin(P1, b), te = part_begin(P1, b + 1);
    const int lane = (int)threadIdx.x % kTX, row = (int)threadIdx.x / kTX;
    double acc[kMaxRhs][4] = {};
    int cur = -1;
    unsigned mask = 0, dmask = 0;
    BImg I{};
    int wx0 = -1, wy0 = 0, top = 0;  // window in LDS, as in kb_step
    for (int t = tb; t < te; ++t) {
        const int im = P1.img[t];
        if (im != cur) {
            if (cur >= 0) flush_slot<4>(acc, slot3 + ((size_t)b * P1.K + (cur - P1.wimg[b])) * 4 * kMaxRhs);
            I = imgs[im];
            mask = dmask = 0;
            for (int j = 0; j < I.nsys; ++j) {
                const Scalars* s = &sc[I.sys0 + j];
                if (tr_selected(s, op)) mask |= 1u << j;
                if (op == 1 && s->verify == 0) dmask |= 1u << j;  // periodic replacement: measure the drift
            }
            cur = im;
            wx0 = -1;
        }
        if (mask == 0) continue;
        const int lt = t - I.t0;
        const int x0 = (lt / I.nty) * kTX, y0 = (lt % I.nty) * kTY;
        const int x = x0 + lane, y = y0 + row;
        const bool valid = x < I.W && y < I.H;
        const int i = y * I.W + x;
        const bool carry = x0 == wx0 && y0 == wy0 + kTY;
        top = carry ? ring_row(top + kTY) : 0;
        PixBands<kF32> pb;
        if (valid) load_pixel_bands<kF32>(I, i, pb);
        auto cell = [&](int j, unsigned o) { return at_off(sys[I.sys0 + j].x, o); };
        float sv[kMaxRhs][kNewPer];
        window_new_load(I, mask, x0, y0, cell, sv);
        window_store(I, mask, x0, y0, carry, top, cell, sv, win);
        wx0 = x0;
        wy0 = y0;
        if (!valid) continue;
        float a[24];
        pixel_bands<kF32>(pb, a);
        int rb[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) rb[q] = ring_row(top + row + q) * kWX + lane;
#pragma unroll
        for (int j = 0; j < kMaxRhs; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const BSys S = sys[I.sys0 + j];
            const float* __restrict__ w = win[j];
            double s = 0.0;
#pragma unroll
            for (int k = 0, kk = 0; k < 25; ++k) {
                if (k == 12) continue;
                s = fma((double)a[kk], (double)w[rb[k / 5] + (k % 5)], s);
                ++kk;
            }
            const unsigned o = 4u * (unsigned)i;
            const float bi = at_off(S.b, o);
            const double tq = fma((double)at_off(S.d, o), (double)w[rb[2] + 2], s) - (double)bi;
            const float out = (float)tq;
            const float zo = at_off(S.dinv, o) * out;
            acc[j][0] += (double)(out * zo);
            acc[j][1] += tq * tq;
            if ((dmask >> j) & 1u) {
                const double dd = (double)out - (double)at_off(S.r, o);
                acc[j][2] += dd * dd;
            }
            acc[j][3] += (double)bi * (double)bi;
            if (op != 0) {
                store_off(S.r, o, out);
                store_off(S.z, o, zo);
            }
        }
    }
    if (cur >= 0) flush_slot<4>(acc, slot3 + ((size_t)b * P1.K + (cur - P1.wimg[b])) * 4 * kMaxRhs);
}

// Finish a true-residual pass: one workgroup per image folds its slots and
// updates the selected systems' scalars.
__global__ __launch_bounds__(kBlock) void kb_trueres_fin(const BImg* __restrict__ imgs, Scalars* sc, BPart P1,
                                                         const double* __restrict__ slot3, int op, double tol) {
    const int im = blockIdx.x;
    const BImg I = imgs[im];
    for (int j = 0; j < I.nsys; ++j) {
        Scalars* s = &sc[I.sys0 + j];
        if (!tr_selected(s, op)) continue;
        const double rz = img_fold(slot3, P1, im, I.wf1, I.wl1, 4 * kMaxRhs, 4 * j);
        const double rr = img_fold(slot3, P1, im, I.wf1, I.wl1, 4 * kMaxRhs, 4 * j + 1);
        const double d2 = img_fold(slot3, P1, im, I.wf1, I.wl1, 4 * kMaxRhs, 4 * j + 2);
        const double bb = img_fold(slot3, P1, im, I.wf1, I.wl1, 4 * kMaxRhs, 4 * j + 3);
        if (threadIdx.x != 0) continue;
        s->g0 = rr;
        if (op == 0) continue;
        if (op == 2) {
            s->bb = bb;
            s->tol2 = tol > 0 ? tol * tol * (bb > 0 ? bb : 1.0) : -1.0;
        }
        replace_residual(s, rz, rr, op == 2, d2, rr);
    }
}

// per system: d = L_ii + lam, dinv = 1/d (0 where d == 0), b = lam*tgt, x = x0 or 0
__global__ __launch_bounds__(kBlock) void kb_set_rhs(const float* __restrict__ Ldiag, int N,
                                                     const float* __restrict__ lam, const float* __restrict__ tgt,
                                                     const float* __restrict__ x0, float* __restrict__ d,
                                                     float* __restrict__ dinv, float* __restrict__ b,
                                                     float* __restrict__ x) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        const float l = lam ? lam[i] : 0.f;
        float di = Ldiag[i];
        float bi = 0.f;
        if (l != 0.f) {
            di += l;
            bi = l * (tgt ? tgt[i] : 0.f);
        }
        d[i] = di;
        dinv[i] = di != 0.f ? 1.f / di : 0.f;
        b[i] = bi;
        x[i] = x0 ? x0[i] : 0.f;
    }
}

// The off-diagonal bands in FP16 or BF16, 8 per pixel in each of three groups: group q
// of pixel i is the 16 bytes at element q * pitch16 + i, halves in band order
// (bands 0..11, 13..24), so a pixel's 24 entries are three aligned 16-byte
// loads (a lane per pixel: 1 KB contiguous per wave and load).
template <int FMT>
__global__ __launch_bounds__(kBlock) void kb_to16(const float* __restrict__ L, size_t pitch, int N,
                                                  unsigned short* __restrict__ L16, size_t pitch16) {
    auto cvt = [](float v) { return (unsigned)(FMT == kBF16 ? f32_to_bf16_rne(v) : f32_to_f16_rne(v)); };
    u32x4* g = reinterpret_cast<u32x4*>(L16);
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            u32x4 v;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const int kk0 = 8 * q + 2 * w, kk1 = kk0 + 1;  // off-diagonal band indices
                const int k0 = kk0 < 12 ? kk0 : kk0 + 1, k1 = kk1 < 12 ? kk1 : kk1 + 1;
                v[w] = cvt(L[(size_t)k0 * pitch + i]) | (cvt(L[(size_t)k1 * pitch + i]) << 16);
            }
            g[(size_t)q * pitch16 + i] = v;
        }
    }
}

// The off-diagonal FP32 bands 4 per pixel in six groups: group q of pixel i is
// the 16 bytes at element q * pitch + i (f32x4 units), bands in order (0..11,
// 13..24), so a pixel's 24 entries are six aligned 16-byte loads.
__global__ __launch_bounds__(kBlock) void kb_group32(const float* __restrict__ L, size_t pitch, int N,
                                                     float* __restrict__ Lg) {
    f32x4* g = reinterpret_cast<f32x4*>(Lg);
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            f32x4 v;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int kk = 4 * q + h, k = kk < 12 ? kk : kk + 1;
                v[h] = L[(size_t)k * pitch + i];
            }
            g[(size_t)q * pitch + i] = v;
        }
    }
}

__global__ void kb_init(Scalars* sc, int nsys, int maxit) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsys) return;
    Scalars z = {};
    z.maxit = maxit;
    z.tol2 = -1.0;
    z.hand.mode = kHandIdle;
    sc[j] = z;
}

}  // namespace gm

using namespace gm;

// Host side of a work-list partition.
struct HostPart {
    int T = 0, G = 0, K = 1;
    std::vector<int> img;   // item -> image
    std::vector<int> wimg;  // workgroup -> image of its first item
    int* dimg = nullptr;
    int* dwimg = nullptr;
    BPart dev() const { return BPart{T, G, K, dimg, dwimg}; }
};

struct gm_batch_s {
    int nimg = 0, nrhs = 0, nsys = 0, precision = 0;
    std::vector<BImg> himg;
    std::vector<BSys> hsys;
    BImg* dimg = nullptr;
    BSys* dsys = nullptr;
    Scalars* dsc = nullptr;
    Scalars* hsc = nullptr;  // pinned, nsys
    HostPart p1, p2;          // step tiles, update chunks
    double* dslot = nullptr;  // slots: step (1 per system), update (2), true residual (4)
    size_t slot1_n = 0, slot2_n = 0, slot3_n = 0;
    float* Lmem = nullptr;
    unsigned short* L16mem = nullptr;
    float* Lgmem = nullptr;  // grouped FP32 bands (FP32 batches)
    float* vmem = nullptr;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_iters = 0, graph_verify = -1;
    hipEvent_t ev = nullptr;
    int per_cu = 16;  // workgroups per CU of the batched launches (gm_batch_set_grid)
};

static void free_part(HostPart& p) {
    if (p.dimg) (void)hipFree(p.dimg);
    if (p.dwimg) (void)hipFree(p.dwimg);
    p.dimg = p.dwimg = nullptr;
}

static void batch_free(gm_batch* b) {
    if (!b) return;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->ev) (void)hipEventDestroy(b->ev);
    free_part(b->p1);
    free_part(b->p2);
    for (void* p : {(void*)b->dimg, (void*)b->dsys, (void*)b->dsc, (void*)b->dslot, (void*)b->Lmem,
                    (void*)b->L16mem, (void*)b->Lgmem, (void*)b->vmem})
        if (p) (void)hipFree(p);
    if (b->hsc) (void)hipHostFree(b->hsc);
    delete b;
}

// Split list p (p.img filled, image-sorted) over G workgroups in contiguous
// ranges; record each image's first and last covering workgroup.
static void split(HostPart& p, int G, std::vector<BImg>& imgs, bool chunks) {
    p.T = (int)p.img.size();
    p.G = G;
    p.wimg.assign(G, 0);
    p.K = 1;
    for (auto& I : imgs) {
        (chunks ? I.wf2 : I.wf1) = G;
        (chunks ? I.wl2 : I.wl1) = -1;
    }
    for (int b = 0; b < G; ++b) {
        const long tb = (long)b * p.T / G, te = (long)(b + 1) * p.T / G;
        if (tb >= te) {
            p.wimg[b] = p.img[std::min<long>(tb, p.T - 1)];
            continue;
        }
        const int i0 = p.img[tb], i1 = p.img[te - 1];
        p.wimg[b] = i0;
        p.K = std::max(p.K, i1 - i0 + 1);
        for (int im = i0; im <= i1; ++im) {
            BImg& I = imgs[im];
            int& wf = chunks ? I.wf2 : I.wf1;
            int& wl = chunks ? I.wl2 : I.wl1;
            wf = std::min(wf, b);
            wl = std::max(wl, b);
        }
    }
}

// Workgroups of a batched launch: per_cu per CU (more than fit at once: a
// workgroup whose range lies in converged images ends at once and the
// hardware starts the next, which balances the uneven work), at most one per
// item, a multiple of the XCD count.
static int batch_grid(const DevInfo& d, long items, int per_cu) {
    long g = std::min((long)d.cus * per_cu, items);
    g = std::max(g, 1L);
    return (int)((g + kNumXcd - 1) / kNumXcd * kNumXcd);
}

// (Re)build both partitions for the batch's grid and upload them.
static gm_status batch_partition(gm_batch* b) {
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    const int G1 = batch_grid(d, (long)b->p1.img.size(), b->per_cu);
    const int G2 = batch_grid(d, (long)b->p2.img.size(), 2 * b->per_cu);
    split(b->p1, G1, b->himg, false);
    split(b->p2, G2, b->himg, true);
    for (HostPart* p : {&b->p1, &b->p2}) {
        free_part(*p);
        GM_HIP(hipMalloc(&p->dimg, sizeof(int) * p->T));
        GM_HIP(hipMalloc(&p->dwimg, sizeof(int) * p->G));
        GM_HIP(hipMemcpy(p->dimg, p->img.data(), sizeof(int) * p->T, hipMemcpyHostToDevice));
        GM_HIP(hipMemcpy(p->dwimg, p->wimg.data(), sizeof(int) * p->G, hipMemcpyHostToDevice));
    }
    b->slot1_n = (size_t)b->p1.G * b->p1.K * kMaxRhs;
    b->slot2_n = (size_t)b->p2.G * b->p2.K * kMaxRhs * 2;
    b->slot3_n = (size_t)b->p1.G * b->p1.K * kMaxRhs * 4;
    if (b->dslot) (void)hipFree(b->dslot);
    b->dslot = nullptr;
    const size_t n = b->slot1_n + b->slot2_n + b->slot3_n;
    GM_HIP(hipMalloc(&b->dslot, sizeof(double) * n));
    GM_HIP(hipMemset(b->dslot, 0, sizeof(double) * n));
    GM_HIP(hipMemcpy(b->dimg, b->himg.data(), sizeof(BImg) * b->nimg, hipMemcpyHostToDevice));
    // a captured iteration graph bakes in the partition
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    return GM_OK;
}

extern "C" {

gm_status gm_batch_create(gm_batch** out, int nimages, const int* widths, const int* heights, int nrhs, int precision) {
    if (!out || nimages <= 0 || !widths || !heights || nrhs <= 0 || nrhs > kMaxRhs)
        return fail(GM_ERR_INVALID, "gm_batch_create: bad arguments (nrhs must be 1..%d)", kMaxRhs);
    if (precision != GM_PREC_FP32 && precision != GM_PREC_FP16 && precision != GM_PREC_BF16)
        return fail(GM_ERR_UNSUPPORTED, "gm_batch_create: batched forms are FP32 and FP16");
    gm_batch* b = new gm_batch();
    b->nimg = nimages;
    b->nrhs = nrhs;
    b->nsys = nimages * nrhs;
    b->precision = precision;
    size_t Lfloats = 0, L16n = 0, Lgn = 0, vfloats = 0;
    std::vector<size_t> loff(nimages), l16off(nimages), lgoff(nimages), voff(nimages);
    for (int i = 0; i < nimages; ++i) {
        const int W = widths[i], H = heights[i];
        // below 2^28 pixels: the grouped band loads address 16 bytes per pixel
        // with a 32-bit byte offset
        if (W < 1 || H < 1 || (long long)W * H >= (1LL << 28)) {
            batch_free(b);
            return fail(GM_ERR_INVALID, "gm_batch_create: image %d is %dx%d (1 to 2^28-1 pixels)", i, W, H);
        }
        const int N = W * H;
        const size_t pitch = (size_t)(N + 63) / 64 * 64;
        BImg I{};
        I.W = W;
        I.H = H;
        I.N = N;
        I.sys0 = i * nrhs;
        I.nsys = nrhs;
        I.nty = (H + kTY - 1) / kTY;
        I.nt = I.nty * ((W + kTX - 1) / kTX);
        I.t0 = (int)b->p1.img.size();
        I.nc = (N + kChunk - 1) / kChunk;
        I.c0 = (int)b->p2.img.size();
        b->p1.img.insert(b->p1.img.end(), I.nt, i);
        b->p2.img.insert(b->p2.img.end(), I.nc, i);
        I.pitch = pitch;
        I.pitch16 = pitch;
        loff[i] = Lfloats;
        Lfloats += 25 * pitch;
        l16off[i] = L16n;
        if (precision != GM_PREC_FP32) L16n += 25 * pitch;
        lgoff[i] = Lgn;
        if (precision == GM_PREC_FP32) Lgn += 24 * pitch;
        voff[i] = vfloats;
        vfloats += (size_t)nrhs * 9 * pitch;
        b->himg.push_back(I);
    }
    hipError_t e = hipSuccess;
#define BTRY(x)                                         \
    do {                                                \
        e = (x);                                        \
        if (e != hipSuccess) {                          \
            batch_free(b);                              \
            return hip_fail(e, #x, __FILE__, __LINE__); \
        }                                               \
    } while (0)
    BTRY(hipMalloc(&b->Lmem, sizeof(float) * Lfloats + 256));
    BTRY(hipMemset(b->Lmem, 0, sizeof(float) * Lfloats + 256));
    if (L16n) {
        BTRY(hipMalloc(&b->L16mem, sizeof(unsigned short) * L16n + 256));
        BTRY(hipMemset(b->L16mem, 0, sizeof(unsigned short) * L16n + 256));
    }
    if (Lgn) {
        BTRY(hipMalloc(&b->Lgmem, sizeof(float) * Lgn + 256));
        BTRY(hipMemset(b->Lgmem, 0, sizeof(float) * Lgn + 256));
    }
    BTRY(hipMalloc(&b->vmem, sizeof(float) * vfloats + 256));
    BTRY(hipMemset(b->vmem, 0, sizeof(float) * vfloats + 256));
    for (int i = 0; i < nimages; ++i) {
        BImg& I = b->himg[i];
        I.L = b->Lmem + loff[i];
        I.L16 = L16n ? b->L16mem + l16off[i] : nullptr;
        I.Lg = Lgn ? b->Lgmem + lgoff[i] : nullptr;
        for (int j = 0; j < nrhs; ++j) {
            float* base = b->vmem + voff[i] + (size_t)j * 9 * I.pitch;
            BSys S;
            S.x = base;
            S.r = base + 1 * I.pitch;
            S.p0 = base + 2 * I.pitch;
            S.p1 = base + 3 * I.pitch;
            S.Ap = base + 4 * I.pitch;
            S.d = base + 5 * I.pitch;
            S.dinv = base + 6 * I.pitch;
            S.b = base + 7 * I.pitch;
            S.z = base + 8 * I.pitch;
            b->hsys.push_back(S);
        }
    }
    BTRY(hipMalloc(&b->dimg, sizeof(BImg) * nimages));
    BTRY(hipMalloc(&b->dsys, sizeof(BSys) * b->nsys));
    BTRY(hipMalloc(&b->dsc, sizeof(Scalars) * b->nsys));
    BTRY(hipMemset(b->dsc, 0, sizeof(Scalars) * b->nsys));
    BTRY(hipMemcpy(b->dsys, b->hsys.data(), sizeof(BSys) * b->nsys, hipMemcpyHostToDevice));
    BTRY(hipHostMalloc(&b->hsc, sizeof(Scalars) * b->nsys, hipHostMallocDefault));
    BTRY(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    BTRY(hipEventCreateWithFlags(&b->ev, hipEventDisableTiming));
#undef BTRY
    gm_status st = batch_partition(b);
    if (st != GM_OK) {
        batch_free(b);
        return st;
    }
    *out = b;
    return GM_OK;
}

gm_status gm_batch_destroy(gm_batch* b) {
    batch_free(b);
    return GM_OK;
}

gm_status gm_batch_assemble(gm_batch* b, int image, const float* im, int nch, size_t im_pitch, float eps,
                            gm_stream_t s) {
    if (!b || image < 0 || image >= b->nimg) return fail(GM_ERR_INVALID, "gm_batch_assemble: bad image index");
    const BImg& I = b->himg[image];
    gm_status st = levin_assemble_raw(const_cast<float*>(I.L), I.pitch, nullptr, nullptr, im, nch, im_pitch, nullptr,
                                      nullptr, nullptr, 0.f, eps, GM_ASSEMBLY_ACCURATE, I.W, I.H, s);
    if (st != GM_OK) return st;
    if (b->precision == GM_PREC_FP32) {
        DevInfo d;
        if ((st = dev_info(&d)) != GM_OK) return st;
        hipLaunchKernelGGL(kb_group32, dim3(elementwise_grid(d, I.N, 1)), dim3(kBlock), 0, (hipStream_t)s, I.L,
                           I.pitch, I.N, const_cast<float*>(I.Lg));
        GM_LAUNCH_CHECK();
    } else {
        DevInfo d;
        if ((st = dev_info(&d)) != GM_OK) return st;
        if (b->precision == GM_PREC_BF16)
            hipLaunchKernelGGL(kb_to16<kBF16>, dim3(elementwise_grid(d, I.N, 1)), dim3(kBlock), 0, (hipStream_t)s,
                               I.L, I.pitch, I.N, const_cast<unsigned short*>(I.L16), I.pitch16);
        else
            hipLaunchKernelGGL(kb_to16<kF16>, dim3(elementwise_grid(d, I.N, 1)), dim3(kBlock), 0, (hipStream_t)s,
                               I.L, I.pitch, I.N, const_cast<unsigned short*>(I.L16), I.pitch16);
        GM_LAUNCH_CHECK();
    }
    return GM_OK;
}

gm_status gm_batch_set_rhs(gm_batch* b, int image, int rhs, const float* lam, const float* tgt, const float* x0,
                           gm_stream_t s) {
    if (!b || image < 0 || image >= b->nimg || rhs < 0 || rhs >= b->nrhs)
        return fail(GM_ERR_INVALID, "gm_batch_set_rhs: bad index");
    const BImg& I = b->himg[image];
    const BSys& S = b->hsys[(size_t)image * b->nrhs + rhs];
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    hipLaunchKernelGGL(kb_set_rhs, dim3(elementwise_grid(d, I.N, 1)), dim3(kBlock), 0, (hipStream_t)s,
                       I.L + 12 * I.pitch, I.N, lam, tgt, x0, S.d, S.dinv, S.b, S.x);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

gm_status gm_batch_alpha(gm_batch* b, int image, int rhs, float** alpha) {
    if (!b || !alpha || image < 0 || image >= b->nimg || rhs < 0 || rhs >= b->nrhs)
        return fail(GM_ERR_INVALID, "gm_batch_alpha: bad index");
    *alpha = b->hsys[(size_t)image * b->nrhs + rhs].x;
    return GM_OK;
}

gm_status gm_batch_info(gm_batch* b, int image, int rhs, gm_solve_info* info) {
    if (!b || !info || image < 0 || image >= b->nimg || rhs < 0 || rhs >= b->nrhs)
        return fail(GM_ERR_INVALID, "gm_batch_info: bad index");
    const Scalars& h = b->hsc[(size_t)image * b->nrhs + rhs];
    const double bb = h.bb > 0 ? h.bb : 1.0;
    info->iterations = h.iters;
    info->status = h.done == kConverged ? GM_OK
                 : h.done == kBreakdown ? GM_ERR_BREAKDOWN
                 : h.done == kNonFinite ? GM_ERR_NONFINITE : GM_ERR_NOT_CONVERGED;
    info->relres = sqrt(h.rr / bb);
    info->true_relres = h.g0 >= 0 ? sqrt(h.g0 / bb) : -1;
    info->replacements = h.replacements;
    return GM_OK;
}

}  // extern "C"

static double* slot1(gm_batch* b) { return b->dslot; }
static double* slot2(gm_batch* b) { return b->dslot + b->slot1_n; }
static double* slot3(gm_batch* b) { return b->dslot + b->slot1_n + b->slot2_n; }

static gm_status batch_launch_iter(gm_batch* b, int verify, hipStream_t hs) {
    const BPart P1 = b->p1.dev(), P2 = b->p2.dev();
    if (b->precision == GM_PREC_FP16)
        hipLaunchKernelGGL(kb_step<kF16>, dim3(P1.G), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, P1, P2, slot1(b),
                           slot2(b), verify);
    else if (b->precision == GM_PREC_BF16)
        hipLaunchKernelGGL(kb_step<kBF16>, dim3(P1.G), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, P1, P2,
                           slot1(b), slot2(b), verify);
    else
        hipLaunchKernelGGL(kb_step<kF32G>, dim3(P1.G), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, P1, P2,
                           slot1(b), slot2(b), verify);
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_update, dim3(P2.G), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, P1, P2, slot1(b),
                       slot2(b));
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status batch_trueres(gm_batch* b, int op, double tol, hipStream_t hs) {
    const BPart P1 = b->p1.dev();
    hipLaunchKernelGGL(kb_trueres, dim3(P1.G), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, P1, op, slot3(b));
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_trueres_fin, dim3(b->nimg), dim3(kBlock), 0, hs, b->dimg, b->dsc, P1, slot3(b), op, tol);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status batch_fold(gm_batch* b, int verify, hipStream_t hs) {
    hipLaunchKernelGGL(kb_fold, dim3(b->nimg), dim3(kBlock), 0, hs, b->dimg, b->dsc, b->p2.dev(), slot2(b), verify);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

// Capture `iters` iterations (step, update pairs) into a graph on the batch's
// own stream (capture needs a non-null stream); replayed on the caller's.
static gm_status batch_graph(gm_batch* b, int iters, int verify) {
    if (b->gexec && b->graph_iters == iters && b->graph_verify == verify) return GM_OK;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    GM_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    gm_status st = GM_OK;
    for (int k = 0; k < iters && st == GM_OK; ++k) st = batch_launch_iter(b, verify, b->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (st != GM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    b->graph = g;
    GM_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->graph_iters = iters;
    b->graph_verify = verify;
    return GM_OK;
}

extern "C" {

gm_status gm_batch_solve(gm_batch* b, const gm_solve_opts* opts, gm_stream_t s) {
    if (!b) return fail(GM_ERR_INVALID, "gm_batch_solve: null batch");
    gm_solve_opts o;
    if (opts) o = *opts;
    else gm_solve_opts_default(&o);
    if (o.method != GM_METHOD_PCG) return fail(GM_ERR_UNSUPPORTED, "gm_batch_solve: batched solver is Jacobi-PCG");
    if (o.max_iter <= 0) return fail(GM_ERR_INVALID, "gm_batch_solve: max_iter");
    gm_status st = GM_OK;
    hipStream_t hs = (hipStream_t)s;
    const int verify = (o.tol > 0 && o.verify) ? 1 : 0;
    const int chunk = o.check_every > 0 ? o.check_every : 10;
    const bool mixed = b->precision != GM_PREC_FP32;
    int refine = o.refine_every > 0 ? o.refine_every : (o.refine_every == 0 && mixed ? 50 : 0);
    if (refine > 0 && refine % chunk) refine = (refine / chunk + 1) * chunk;  // replacement at chunk boundaries

    hipLaunchKernelGGL(kb_init, dim3((b->nsys + 63) / 64), dim3(64), 0, hs, b->dsc, b->nsys, o.max_iter);
    GM_LAUNCH_CHECK();
    if ((st = batch_trueres(b, 2, o.tol, hs)) != GM_OK) return st;  // r = A x - b; ||b||^2 -> tol2
    if ((st = batch_graph(b, chunk, verify)) != GM_OK) return st;

    long enq = 0;
    bool finished = false;
    while (!finished) {
        GM_HIP(hipGraphLaunch(b->gexec, hs));
        enq += chunk;
        if (refine > 0 && enq % refine == 0) {
            if ((st = batch_trueres(b, 1, o.tol, hs)) != GM_OK) return st;
        }
        GM_HIP(hipMemcpyAsync(b->hsc, b->dsc, sizeof(Scalars) * b->nsys, hipMemcpyDeviceToHost, hs));
        GM_HIP(hipEventRecord(b->ev, hs));
        GM_HIP(hipEventSynchronize(b->ev));
        int running = 0, verifying = 0;
        for (int j = 0; j < b->nsys; ++j) {
            if (b->hsc[j].done == kRun) {
                ++running;
                if (b->hsc[j].verify) ++verifying;
            }
        }
        if (running == 0 || enq >= o.max_iter) finished = true;
        if (verifying > 0 && !finished) {
            if ((st = batch_trueres(b, 3, o.tol, hs)) != GM_OK) return st;
        }
    }
    // fold the last update's partials, check any recursive convergence it
    // shows, then the FP64 true residual of every system (g0) and a host copy
    if ((st = batch_fold(b, verify, hs)) != GM_OK) return st;
    if ((st = batch_trueres(b, 3, o.tol, hs)) != GM_OK) return st;
    if ((st = batch_trueres(b, 0, o.tol, hs)) != GM_OK) return st;
    GM_HIP(hipMemcpyAsync(b->hsc, b->dsc, sizeof(Scalars) * b->nsys, hipMemcpyDeviceToHost, hs));
    GM_HIP(hipStreamSynchronize(hs));
    int bad = 0;
    for (int j = 0; j < b->nsys; ++j)
        if (b->hsc[j].done != kConverged) ++bad;
    if (bad && o.tol > 0) return fail(GM_ERR_NOT_CONVERGED, "gm_batch_solve: %d of %d systems did not converge", bad,
                                      b->nsys);
    return GM_OK;
}

gm_status gm_batch_set_grid(gm_batch* b, int per_cu) {
    if (!b || per_cu < 1 || per_cu > 64) return fail(GM_ERR_INVALID, "gm_batch_set_grid: per_cu 1..64");
    b->per_cu = per_cu;
    return batch_partition(b);
}

int gm_batch_tiles(gm_batch* b) { return b ? b->p1.T : 0; }

}  // extern "C"
