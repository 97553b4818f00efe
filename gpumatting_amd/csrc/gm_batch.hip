// gm_batch.hip -- many images of uneven size, several right-hand sides each,
// solved together (no reference counterpart).
//
// Each image i has one Levin Laplacian L_i (25 FP32 bands, no data term) and
// nrhs independent constraint sets j, i.e. nrhs systems
//     (L_i + diag(lam_ij)) x_ij = lam_ij * t_ij .
// The systems of one image share L_i's off-diagonal bands: a pass over a tile
// loads those 24 bands once and applies them to every right-hand side; each
// system brings its own FP32 diagonal d_ij = L_ii + lam_ij.  Jacobi-PCG per
// system (z = D^-1 r), one launch per CG iteration for all systems.
//
// The fused iteration.  Launch k takes every running system from state j
// (x_j, r_j, p_j, Ap_j and the five sums of that state) to state j + 1:
//   alpha_j = <r_j, z_j> / <p_j, Ap_j>,   x_{j+1} = x_j + alpha_j p_j,
//   r_{j+1} = r_j + alpha_j Ap_j,
//   beta_{j+1} = (<r_j,z_j> + 2 alpha_j <Ap_j,z_j> + alpha_j^2 <Ap_j,D^-1 Ap_j>) / <r_j,z_j>
//                (the square <r_{j+1}, D^-1 r_{j+1}> expanded, so beta is known
//                before any vector is read),
//   p_{j+1} = -z_{j+1} + beta_{j+1} p_j,   Ap_{j+1} = A p_{j+1},
// and leaves the sums of state j + 1: <p,Ap>, <r,z>, ||r||^2 (exact, of the
// r it stores), <Ap,z>, <Ap,D^-1 Ap>.  The launch starts by folding the sums
// the previous launch left (per image, fixed order), which also gives the
// stopping test on the exact ||r_j||.  r, p and Ap alternate between two
// buffers by iteration parity (neighbouring tiles read r_j, Ap_j, p_j at their
// window cells while the owners write the new ones); x is updated in place.
// A system whose residual was just formed from scratch (start, replacement,
// restart) takes a "first" step instead: p_j = -z_j + beta p_{j-1}, Ap_j, the
// sums, and no update.
//
// Tiles.  TX x TY pixels (TX = 60, 124 or 252: windows 64, 128 or 256 wide, one
// new window cell per thread), numbered down each image's TX-wide columns in
// segments of a range's worth of tile rows (bplan::tile_origin); a tile's
// window (2-pixel halo) shares its top
// 4 rows with the tile above and is an LDS ring, so a workgroup that has just
// done the tile above loads only the TY new rows.  A tile advances r and x at
// the cells of its new rows (bplan::advances), so going down a column every
// pixel is advanced exactly once.  A launch splits the tile list over its
// workgroups in contiguous ranges (gm_batch_plan.h); a workgroup keeps its
// partial sums per (image, system) in registers and stores one slot per image
// it touches; the next launch folds an image's slots over the ranges covering
// it.  No fence or atomic: the kernel boundary orders the stores.  Chunks of
// iterations are captured in hipGraphs (kernel nodes only).  A workgroup reads
// each system's scalar state once (thread 0) and shares it through LDS, so its
// control flow is uniform even while another workgroup publishes a stop.
//
// Mixed precision: the off-diagonal bands are also kept in FP16 or BF16
// (unscaled; |L_ij| < 20), the per-system diagonals stay FP32, and the
// residual is replaced by the FP64-accumulated true residual (FP32 L) every
// 50 iterations (see kDrift2 for when the search direction is kept).
#include <string.h>

#include <algorithm>
#include <vector>

#include "gm_batch_plan.h"
#include "gm_internal.h"

namespace gm {

gm_status levin_assemble_raw(float* A, size_t apitch, float* b, float* dinv, const float* im, int nch, size_t pitch,
                             const float* lam, const float* tgt, const unsigned char* scr, float gamma, float eps,
                             int mode, int W, int H, gm_stream_t s);

using bplan::part_begin;
using bplan::range_of;
using bplan::ring_row;
using bplan::slot_index;

constexpr int kMaxRhs = 8;
constexpr int kBatchMaxGrid = 16384;  // workgroups of a batched launch, at most
constexpr int kNQ = 5;                           // sums per system of a fused launch: <p,Ap>, <r,z>, ||r||^2, <Ap,z>, <Ap,D^-1 Ap>
constexpr int kRTX = 128, kRTY = kBlock / kRTX;  // true-residual tiles: 128 x 2
constexpr int kChunk = 4 * kBlock;               // update-only chunks: 1024 pixels, 4 per thread
constexpr int kKeep = 4;                         // window rows a tile shares with the tile directly above

// Fused-step tiles of window width WX (64, 128, 256): TX = WX - 4 pixels wide,
// TY = kBlock / WX rows, so the TY new window rows a tile loads are exactly
// one cell per thread (loaded before the previous tile is done with the
// window).  Thread t is window column t % WX of tile row t / WX; the last 4
// columns' threads have no pixel of their own.
template <int WX_>
struct BGeom {
    static constexpr int WX = WX_, TX = WX - 4, TY = kBlock / WX, WY = TY + 4;
    static constexpr int NEW = (WY - kKeep) * WX;  // cells of the window rows below the shared ones
    static constexpr int KEEP = kKeep * WX;
    static_assert(TY * WX == kBlock && TY <= kKeep && NEW == kBlock, "one new window cell per thread");
};

// Device view of one partition.
struct BPart {
    int T, G, K;
    const int* img;   // item -> image
    const int* wimg;  // range -> image of its first item
};

struct BImg {
    int W, H, N, sys0, nsys;
    int t0, nty, segf;  // fused tiles: first index in the fused list, tiles per image column, segment length
    int q0, nqy, segq;  // true-residual tiles (128 x 2)
    int c0;        // update-only chunks
    int wf1, wl1;  // ranges of the fused partition covering the image
    int wf3, wl3;  // ... of the true-residual partition
    int wf2;       // first range of the chunk partition covering the image
    const float* L;
    size_t pitch;
    const unsigned short* L16;
    const float* Lg;  // FP32 form: the off-diagonal bands 4 per pixel in 6 groups (kb_group32)
    float* V;         // the systems' vectors: system j's block of kSysVecs vectors of pitch floats at V + j * kSysVecs * pitch
};

struct BSys {
    float *x, *r[2], *p[2], *Ap[2], *d, *b;
};
// A system's block: x, r[0], r[1], p[0], p[1], Ap[0], Ap[1], d, b in that order,
// pitch floats each (BImg::V).  Images have fewer than 2^25 pixels, so a
// block spans less than 2^31 bytes and one range-checked buffer descriptor
// with 32-bit offsets covers it.
constexpr int kSysVecs = 9;
enum { kVx = 0, kVr = 1, kVp = 3, kVAp = 5, kVd = 7, kVb = 8 };

// ---------------------------------------------------------------------------
// slots and folds

// Sum an image's slots over the ranges r in [wf, wl] (fixed order): NQ
// quantities at slot[slot_index(r, ...) + q]; every thread gets the totals.
template <int NQ>
__device__ __forceinline__ void img_fold(const double* __restrict__ slot, const BPart& P, int im, int wf, int wl,
                                         int stride, int off, double (&out)[NQ]) {
    __shared__ double fr[NQ][kWaves];
    __shared__ double ft[NQ];
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) s[q] = 0;
    for (int r = wf + (int)threadIdx.x; r <= wl; r += kBlock) {
        const double* src = slot + slot_index(r, P.K, im, P.wimg[r], stride, off);
#pragma unroll
        for (int q = 0; q < NQ; ++q) s[q] += src[q];
    }
    __syncthreads();  // an earlier call's readers are done with fr / ft
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double t = wave_sum(s[q]);
        if ((threadIdx.x & 63) == 0) fr[q][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += fr[threadIdx.x][w];
        ft[threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = ft[q];
}

// Store this workgroup's per-system partial sums for one image (block
// reduction of acc[j][q]) at dst[j * NQ + q], and clear them.
template <int NR, int NQ>
__device__ __forceinline__ void flush_slot(double (&acc)[NR][NQ], double* __restrict__ dst) {
    __shared__ double red[NR * NQ][kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double s = wave_sum(acc[j][q]);
            if (lane == 0) red[j * NQ + q][wid] = s;
            acc[j][q] = 0.0;
        }
    __syncthreads();
    if (threadIdx.x < NR * NQ) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += red[threadIdx.x][w];
        dst[threadIdx.x] = t;
    }
    __syncthreads();  // red is reused by the next flush
}

// ---------------------------------------------------------------------------
// vector and band access

// Element at a byte offset from a vector's base, in global memory (the
// pointers come from device tables, so the compiler cannot tell by itself and
// would use flat accesses, which also count against the LDS wait counter).
// Images have fewer than 2^28 pixels, so byte offsets of up to 16 bytes per
// pixel fit 32 bits.
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

// the FP32 bands in groups of 4 per pixel (kb_group32)
constexpr int kF32G = 3;
// The 24 off-diagonal band values of pixel i (band order, 12 skipped), as
// loaded: FP32 in six 16-byte groups, or the 16-bit bands in three.
template <int FMT>
struct PixBands {
    u32x4 g[3];
};
template <>
struct PixBands<kF32G> {
    f32x4 g[6];
};
template <int FMT>
__device__ __forceinline__ void load_pixel_bands(const BImg& I, int i, PixBands<FMT>& pb) {
    if constexpr (FMT == kF32G) {
        const f32x4* g = reinterpret_cast<const f32x4*>(I.Lg);
#pragma unroll
        for (int q = 0; q < 6; ++q) pb.g[q] = nt_at_off(g + (size_t)q * I.pitch, 16u * (unsigned)i);
    } else {
        const u32x4* g = reinterpret_cast<const u32x4*>(I.L16);
#pragma unroll
        for (int q = 0; q < 3; ++q) pb.g[q] = nt_at_off(g + (size_t)q * I.pitch, 16u * (unsigned)i);
    }
}
template <int FMT>
__device__ __forceinline__ void pixel_bands(const PixBands<FMT>& pb, float (&a)[24]) {
    if constexpr (FMT == kF32G) {
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

__device__ __forceinline__ float inv_diag(float d) { return d != 0.f ? 1.f / d : 0.f; }

// ---------------------------------------------------------------------------
// the fused iteration

enum { kSysIdle = 0, kSysNormal = 1, kSysFirst = 2 };

// What a workgroup does for one system of the current image.
struct SysStep {
    int mode;
    float alpha, beta;
    int it;  // iterations done (state j)
};

// cst[j] as wave-uniform (scalar) values: the per-system choices and buffer
// pointers then live in scalar registers.
__device__ __forceinline__ SysStep uniform_step(const SysStep& s) {
    SysStep u;
    u.mode = __builtin_amdgcn_readfirstlane(s.mode);
    u.it = __builtin_amdgcn_readfirstlane(s.it);
    u.alpha = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s.alpha)));
    u.beta = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s.beta)));
    return u;
}

// A system's scalar state as one thread read it at image entry.
struct SysEnt {
    int run, fresh, iters, maxit;
    float beta;
    double rz_prev, tol2;
    FHand h;
};

// Thread j < nsys reads system j's state into ent[j] (then a barrier).
template <int NR>
__device__ __forceinline__ void read_entry(const BImg& I, const Scalars* sc, int k, SysEnt* ent) {
    __syncthreads();  // ent / cst of the previous image are no longer read
    const int j = (int)threadIdx.x;
    if (j < NR && j < I.nsys) {
        const Scalars* s = &sc[I.sys0 + j];
        SysEnt e;
        e.run = s->done == kRun && !s->verify;
        e.fresh = s->fresh;
        e.iters = s->iters;
        e.maxit = s->maxit;
        e.beta = s->beta;
        e.rz_prev = s->rz_prev;
        e.tol2 = s->tol2;
        e.h = s->fh[k & 1];
        ent[j] = e;
    }
    __syncthreads();
}

// Stopping test on a folded state (iteration count h.iters).
__device__ __forceinline__ int stop_code(const SysEnt& e, double pAp, double rz, double rr, double q3, double q4,
                                         int verify, int* ver) {
    *ver = 0;
    if (!isfinite(pAp) || !isfinite(rz) || !isfinite(rr) || !isfinite(q3) || !isfinite(q4)) return kNonFinite;
    if (e.tol2 >= 0 && rr <= e.tol2) {
        if (verify) {
            *ver = 1;
            return kRun;
        }
        return kConverged;
    }
    if (e.h.iters >= e.maxit) return kMaxIt;
    if (!(pAp > 0)) return kBreakdown;
    return kRun;
}

// Publish a stopped system's state (the image's lead thread).
__device__ __forceinline__ void publish_stop(Scalars* s, const SysEnt& e, double pAp, double rz, double rr, int stop,
                                             int ver) {
    s->rz = rz;
    s->rz_prev = e.h.rr_prev;
    s->rr = rr;
    s->beta = e.h.beta;
    s->pAp = pAp;
    s->iters = e.h.iters;
    s->pend = 0;
    s->verify = ver;
    s->done = stop;
}

// Per-system decisions at image entry (results in cst, every thread): fold
// the previous launch's sums, stopping test, alpha and beta.  The image's
// first covering range publishes stops and writes the next slot of the
// hand-off.  tune: every system steps with alpha = beta = 0.5 from parity 0
// and nothing is read or published.
template <int NR>
__device__ __forceinline__ unsigned fused_enter(const BImg& I, int im, Scalars* sc, const double* __restrict__ sin,
                                                const BPart& P, int k, int tune, bool lead_img, int verify,
                                                SysEnt* ent, SysStep* cst) {
    read_entry<NR>(I, sc, k, ent);
    const bool lead = lead_img && threadIdx.x == 0;
    unsigned mask = 0;
    for (int j = 0; j < I.nsys && j < NR; ++j) {
        const SysEnt e = ent[j];
        Scalars* s = &sc[I.sys0 + j];
        SysStep st{kSysIdle, 0.f, 0.f, 0};
        if (tune) {
            st = SysStep{kSysNormal, 0.5f, 0.5f, 0};
        } else if (e.run) {
            if (k == 0 && e.fresh) {
                st = SysStep{kSysFirst, 0.f, e.beta, e.iters};
                if (lead) s->fh[(k + 1) & 1] = FHand{e.rz_prev, e.beta, e.iters, 0, 0};
            } else {
                double t[kNQ];
                img_fold<kNQ>(sin, P, im, I.wf1, I.wl1, kNQ * NR, kNQ * j, t);
                const double pAp = t[0], rz = t[1], rr = t[2], rzAp = t[3], ApDAp = t[4];
                int ver;
                const int stop = stop_code(e, pAp, rz, rr, rzAp, ApDAp, verify, &ver);
                if (stop != kRun || ver) {
                    if (lead) publish_stop(s, e, pAp, rz, rr, stop, ver);
                } else {
                    const float al = (float)(rz / pAp);
                    const double a = (double)al, rz1 = rz + 2.0 * a * rzAp + a * a * ApDAp;
                    const float be = rz1 > 0 ? (float)(rz1 / rz) : 0.f;
                    st = SysStep{kSysNormal, al, be, e.h.iters};
                    if (lead) s->fh[(k + 1) & 1] = FHand{rz, be, e.h.iters + 1, 0, 0};
                }
            }
        }
        if (threadIdx.x == 0) cst[j] = st;
        if (st.mode != kSysIdle) mask |= 1u << j;
    }
    __syncthreads();
    return mask;
}

// One system's vectors for the step from state `it` in mode `mode`: a
// range-checked buffer over the system's block (a load out of range reads 0,
// a store out of range is dropped) and the byte offsets of the vectors in it,
// all wave-uniform, so a cell's loads need no pointer table and no branch.
struct SysBuf {
    __amdgpu_buffer_rsrc_t rs;
    int rin, rout, apin, apout, pold, pnew, x, d;
};
__device__ __forceinline__ SysBuf sys_buf(const BImg& I, int j, const SysStep& st) {
    const int a = st.it & 1, b = a ^ 1;
    const int P = 4 * (int)I.pitch;  // bytes per vector
    SysBuf q;
    q.rs = __builtin_amdgcn_make_buffer_rsrc(I.V + (size_t)j * kSysVecs * I.pitch, 0, kSysVecs * P, 0x00020000);
    q.x = kVx * P;
    q.rin = (kVr + a) * P;
    q.rout = (kVr + b) * P;
    q.apin = (kVAp + a) * P;
    q.d = kVd * P;
    if (st.mode == kSysFirst) {  // p_j from p_{j-1} (the other buffer), written where p_j lives
        q.pold = (kVp + b) * P;
        q.pnew = (kVp + a) * P;
        q.apout = (kVAp + a) * P;
    } else {
        q.pold = (kVp + a) * P;
        q.pnew = (kVp + b) * P;
        q.apout = (kVAp + b) * P;
    }
    return q;
}
// Byte offset of element g of a vector, or an offset past every block (the
// access then reads 0 or writes nothing).
__device__ __forceinline__ unsigned sb_off(bool in, long long g) { return in ? 4u * (unsigned)g : 0x80000000u; }
__device__ __forceinline__ float sb_ld(const SysBuf& q, unsigned off, int vec) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(q.rs, (int)off, vec, 0));
}
__device__ __forceinline__ void sb_st(const SysBuf& q, unsigned off, int vec, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), q.rs, (int)off, vec, 0);
}

// One staged window cell of one system: r, Ap, p_old, d (and x where the cell
// is advanced here).  in: the cell is a pixel of the image (else all 0).
struct CellIn {
    float r, a, p, d, x;
};

__device__ __forceinline__ CellIn load_cell(const SysBuf& q, const SysStep& st, bool in, bool adv, long long g) {
    CellIn v;
    const unsigned o = sb_off(in, g);
    v.r = sb_ld(q, o, q.rin);
    v.a = st.mode == kSysNormal ? sb_ld(q, o, q.apin) : 0.f;
    v.p = (st.beta != 0.f || st.mode == kSysNormal) ? sb_ld(q, o, q.pold) : 0.f;
    v.d = sb_ld(q, o, q.d);
    v.x = st.mode == kSysNormal ? sb_ld(q, sb_off(adv, g), q.x) : 0.f;  // adv implies in
    return v;
}

template <int FMT, int NR, int WXT>
__global__ __launch_bounds__(kBlock) void kb_fused(const BImg* __restrict__ imgs, Scalars* sc, BPart P,
                                                   double* __restrict__ slots, size_t half, int k, int tune,
                                                   int verify) {
    using Gm = BGeom<WXT>;
    constexpr int TX = Gm::TX, TY = Gm::TY, WX = Gm::WX, WY = Gm::WY;
    __shared__ float wp[NR][WY * WX];  // p_{j+1} of the window cells
    __shared__ f32x2 wzd[NR][WY * WX];  // z_{j+1} and d of the window cells (read at the thread's own pixel)
    __shared__ SysEnt ent[NR];
    __shared__ SysStep cst[NR];
    const int r = range_of(blockIdx.x, P.G);
    const int tb = part_begin(r, P.T, P.G), te = part_begin(r + 1, P.T, P.G);
    const double* __restrict__ sin = slots + (size_t)(k & 1) * half;
    double* __restrict__ sout = slots + (size_t)((k + 1) & 1) * half;
    const int lane = (int)threadIdx.x % WX, row = (int)threadIdx.x / WX;
    double acc[NR][kNQ] = {};
    int cur = -1;
    unsigned mask = 0;
    BImg I{};
    int wx0 = -1, wy0 = 0, top = 0;  // the window in LDS: origin of its tile (none: wx0 < 0), ring row of window row 0
    for (int t = tb; t < te; ++t) {
        const int im = P.img[t];
        if (im != cur) {
            if (cur >= 0) flush_slot<NR, kNQ>(acc, sout + slot_index(r, P.K, cur, P.wimg[r], kNQ * NR, 0));
            I = imgs[im];
            mask = fused_enter<NR>(I, im, sc, sin, P, k, tune, r == I.wf1, verify, ent, cst);
            cur = im;
            wx0 = -1;
        }
        if (mask == 0) continue;
        int x0, y0;
        bplan::tile_origin(t - I.t0, I.nty, (I.W + TX - 1) / TX, I.segf, TX, TY, x0, y0);
        const int x = x0 + lane, y = y0 + row;
        const bool valid = lane < TX && x < I.W && y < I.H;
        const int i = y * I.W + x;
        // the tile directly above is in LDS: its window rows TY.. are this one's rows 0..
        const bool carry = x0 == wx0 && y0 == wy0 + TY;
        top = carry ? ring_row(top + TY, WY) : 0;
        // every global load of the tile first
        PixBands<FMT> pb;
        if (valid) load_pixel_bands<FMT>(I, i, pb);
        CellIn cin[NR];
        // new window cell threadIdx.x
        const int nwy = kKeep + row, nwx = lane;
        const long long ng = (long long)(y0 - 2 + nwy) * I.W + (x0 - 2 + nwx);
        const bool nin = ng >= 0 && ng < I.N;
        const bool nadv = bplan::advances(nwy, nwx, x0, y0, TX, I.W, I.H);
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const SysStep st = uniform_step(cst[j]);
            cin[j] = load_cell(sys_buf(I, j, st), st, nin, nadv, ng);
        }
        __syncthreads();  // the previous tile is done with the window
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const SysStep st = uniform_step(cst[j]);
            const SysBuf q = sys_buf(I, j, st);
            // r_{j+1}, z_{j+1}, p_{j+1} of a cell into the window; x and r
            // advanced where this tile advances the cell (adv implies in)
            auto step_cell = [&](const CellIn& v, bool in, bool adv, long long g, int li) {
                const float rn = st.mode == kSysNormal ? fmaf(st.alpha, v.a, v.r) : v.r;
                float zn = inv_diag(v.d) * rn;
                float pn = st.beta != 0.f ? fmaf(st.beta, v.p, -zn) : -zn;
                if (!in) pn = zn = 0.f;
                if (adv) {
                    acc[j][1] += (double)rn * (double)zn;
                    acc[j][2] += (double)rn * (double)rn;
                }
                if (st.mode == kSysNormal) {
                    const unsigned o = sb_off(adv, g);
                    sb_st(q, o, q.x, fmaf(st.alpha, v.p, v.x));
                    sb_st(q, o, q.rout, rn);
                }
                wp[j][li] = pn;
                wzd[j][li] = f32x2{zn, in ? v.d : 0.f};
            };
            // a kept-row window cell, loaded now (a run's first tile, or the
            // first of a new column)
            auto cell_now = [&](int wy, int wx) {
                const long long g = (long long)(y0 - 2 + wy) * I.W + (x0 - 2 + wx);
                const bool in = g >= 0 && g < I.N;
                const bool adv = bplan::advances(wy, wx, x0, y0, TX, I.W, I.H);
                step_cell(load_cell(q, st, in, adv, g), in, adv, g, ring_row(top + wy, WY) * WX + wx);
            };
            if (!carry)
                for (int e = (int)threadIdx.x; e < Gm::KEEP; e += kBlock) cell_now(e / WX, e % WX);
            step_cell(cin[j], nin, nadv, ng, ring_row(top + nwy, WY) * WX + nwx);
        }
        __syncthreads();
        wx0 = x0;
        wy0 = y0;
        if (!valid) continue;
        float a[24];
        pixel_bands<FMT>(pb, a);
        // LDS offsets of the thread's 5 window rows (cell (row + s, lane + dx) is pixel (y - 2 + s, x - 2 + dx))
        int rb[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) rb[s] = ring_row(top + row + s, WY) * WX + lane;
        const unsigned oi = 4u * (unsigned)i;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const SysBuf q = sys_buf(I, j, uniform_step(cst[j]));
            const float* __restrict__ w = wp[j];
            float yv = 0.f;
#pragma unroll
            for (int kb = 0, kk = 0; kb < 25; ++kb) {
                if (kb == 12) continue;
                yv = fmaf(a[kk], w[rb[kb / 5] + (kb % 5)], yv);
                ++kk;
            }
            const float pc = w[rb[2] + 2];
            const f32x2 zd = wzd[j][rb[2] + 2];
            yv = fmaf(zd.y, pc, yv);
            sb_st(q, oi, q.pnew, pc);
            sb_st(q, oi, q.apout, yv);
            acc[j][0] += (double)(pc * yv);
            acc[j][3] += (double)yv * (double)zd.x;
            acc[j][4] += (double)yv * ((double)yv * (double)inv_diag(zd.y));
        }
    }
    if (cur >= 0) flush_slot<NR, kNQ>(acc, sout + slot_index(r, P.K, cur, P.wimg[r], kNQ * NR, 0));
}

// The update only (the launch before a residual replacement, whose step would
// be redone from the new residual): x_{j+1}, r_{j+1} of every cell, state j + 1
// published with pend = 1 (its sums are the replacement's to form).  Chunks
// of 1024 pixels per workgroup range; launch index k as in kb_fused.
template <int NR>
__global__ __launch_bounds__(kBlock) void kb_fupdate(const BImg* __restrict__ imgs, const BSys* __restrict__ sys,
                                                     Scalars* sc, BPart P1, BPart P2, const double* __restrict__ slots,
                                                     size_t half, int k, int verify) {
    __shared__ SysEnt ent[NR];
    __shared__ SysStep cst[NR];
    const int r = range_of(blockIdx.x, P2.G);
    const int cb = part_begin(r, P2.T, P2.G), ce = part_begin(r + 1, P2.T, P2.G);
    const double* __restrict__ sin = slots + (size_t)(k & 1) * half;
    int cur = -1;
    unsigned mask = 0;
    BImg I{};
    for (int c = cb; c < ce; ++c) {
        const int im = P2.img[c];
        if (im != cur) {
            I = imgs[im];
            read_entry<NR>(I, sc, k, ent);
            const bool lead = r == I.wf2 && threadIdx.x == 0;
            mask = 0;
            for (int j = 0; j < I.nsys && j < NR; ++j) {
                const SysEnt e = ent[j];
                Scalars* s = &sc[I.sys0 + j];
                SysStep st{kSysIdle, 0.f, 0.f, 0};
                if (e.run) {
                    double t[kNQ];
                    img_fold<kNQ>(sin, P1, im, I.wf1, I.wl1, kNQ * NR, kNQ * j, t);
                    const double pAp = t[0], rz = t[1], rr = t[2];
                    int ver;
                    const int stop = stop_code(e, pAp, rz, rr, t[3], t[4], verify, &ver);
                    if (stop != kRun || ver) {
                        if (lead) publish_stop(s, e, pAp, rz, rr, stop, ver);
                    } else {
                        st = SysStep{kSysNormal, (float)(rz / pAp), 0.f, e.h.iters};
                        if (lead) {
                            s->rz = rz;
                            s->rz_prev = e.h.rr_prev;
                            s->rr = rr;
                            s->beta = e.h.beta;
                            s->pAp = pAp;
                            s->iters = e.h.iters + 1;
                            s->pend = 1;
                        }
                    }
                }
                if (threadIdx.x == 0) cst[j] = st;
                if (st.mode != kSysIdle) mask |= 1u << j;
            }
            __syncthreads();
            cur = im;
        }
        if (mask == 0) continue;
        const int i0 = (c - I.c0) * kChunk + 4 * (int)threadIdx.x;
        if (i0 >= I.N) continue;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const SysStep st = uniform_step(cst[j]);
            const BSys& S = sys[I.sys0 + j];
            const int a = st.it & 1;
            const float *p = S.p[a], *ap = S.Ap[a], *rin = S.r[a];
            float* rout = S.r[a ^ 1];
            const float al = st.alpha;
            if (i0 + 4 <= I.N) {  // per-system vectors are 256-byte aligned and i0 % 4 == 0
                const unsigned o = 4u * (unsigned)i0;
                const f32x4 xv = at_off4(S.x, o), rv = at_off4(rin, o), pv = at_off4(p, o), av = at_off4(ap, o);
                f32x4 xn, rn;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    xn[u] = fmaf(al, pv[u], xv[u]);
                    rn[u] = fmaf(al, av[u], rv[u]);
                }
                store_off4(S.x, o, xn);
                store_off4(rout, o, rn);
            } else {
                for (int ii = i0; ii < I.N; ++ii) {
                    const unsigned o = 4u * (unsigned)ii;
                    store_off(S.x, o, fmaf(al, at_off(p, o), at_off(S.x, o)));
                    store_off(rout, o, fmaf(al, at_off(ap, o), at_off(rin, o)));
                }
            }
        }
    }
}

// End of a chunk of fused launches (one workgroup per image): publish every
// running system's state from the sums the last launch left (parity `par`),
// with the stopping test, so that the host sees it; clear the fresh marks
// (the chunk's first launch consumed them).
template <int NR>
__global__ __launch_bounds__(kBlock) void kb_ffold(const BImg* __restrict__ imgs, Scalars* sc, BPart P1,
                                                   const double* __restrict__ slots, size_t half, int par,
                                                   int verify) {
    __shared__ SysEnt ent[NR];
    const int im = blockIdx.x;
    const BImg I = imgs[im];
    const double* __restrict__ sin = slots + (size_t)par * half;
    read_entry<NR>(I, sc, par, ent);
    for (int j = 0; j < I.nsys && j < NR; ++j) {
        const SysEnt e = ent[j];
        Scalars* s = &sc[I.sys0 + j];
        double t[kNQ];
        if (e.run) img_fold<kNQ>(sin, P1, im, I.wf1, I.wl1, kNQ * NR, kNQ * j, t);
        if (threadIdx.x != 0) continue;
        s->fresh = 0;
        if (!e.run) continue;
        const double pAp = t[0], rz = t[1], rr = t[2];
        int ver;
        const int stop = stop_code(e, pAp, rz, rr, t[3], t[4], verify, &ver);
        publish_stop(s, e, pAp, rz, rr, stop == kBreakdown ? kRun : stop, ver);
    }
}

// systems a true-residual pass selects (op: see kb_trueres)
__device__ __forceinline__ bool tr_selected(const Scalars* s, int op) {
    return op == 0 || op == 2 || (op == 1 && s->done == kRun) || (op == 3 && s->done == kRun && s->verify);
}

// The window of the true-residual kernel (128 x 2 tiles, x staged; a ring as
// in kb_fused).
constexpr int kRWX = kRTX + 4, kRWY = kRTY + 4;
constexpr int kRNew = (kRWY - kKeep) * kRWX;
constexpr int kRNewPer = (kRNew + kBlock - 1) / kBlock;

// FP64-accumulated true residual t = (L_off + diag d) x - b of the selected
// systems, partials (<t, D^-1 t>, ||t||^2, drift ||t - r||^2, ||b||^2) per
// image slot; kb_trueres_fin folds them.  The residual lives in the buffer of
// the system's iteration parity.  op 0: measure only (g0); 1: replace r for
// running systems; 2: start (all systems, beta = 0, also ||b||^2); 3: replace
// r for systems whose recursive residual asked for verification.
__global__ __launch_bounds__(kBlock) void kb_trueres(const BImg* __restrict__ imgs, const BSys* __restrict__ sys,
                                                     const Scalars* sc, BPart P3, int op, double* __restrict__ slot3) {
    __shared__ float win[kMaxRhs][kRWY * kRWX];
    const int r = range_of(blockIdx.x, P3.G);
    const int tb = part_begin(r, P3.T, P3.G), te = part_begin(r + 1, P3.T, P3.G);
    const int lane = (int)threadIdx.x % kRTX, row = (int)threadIdx.x / kRTX;
    double acc[kMaxRhs][4] = {};
    int cur = -1;
    unsigned mask = 0, dmask = 0;
    int par[kMaxRhs] = {};
    BImg I{};
    int wx0 = -1, wy0 = 0, top = 0;
    for (int t = tb; t < te; ++t) {
        const int im = P3.img[t];
        if (im != cur) {
            if (cur >= 0) flush_slot<kMaxRhs, 4>(acc, slot3 + slot_index(r, P3.K, cur, P3.wimg[r], 4 * kMaxRhs, 0));
            I = imgs[im];
            mask = dmask = 0;
            for (int j = 0; j < I.nsys; ++j) {
                const Scalars* s = &sc[I.sys0 + j];
                if (tr_selected(s, op)) mask |= 1u << j;
                if (op == 1 && s->verify == 0) dmask |= 1u << j;  // periodic replacement: measure the drift
                par[j] = s->iters & 1;
            }
            cur = im;
            wx0 = -1;
        }
        if (mask == 0) continue;
        int x0, y0;
        bplan::tile_origin(t - I.q0, I.nqy, (I.W + kRTX - 1) / kRTX, I.segq, kRTX, kRTY, x0, y0);
        const int x = x0 + lane, y = y0 + row;
        const bool valid = x < I.W && y < I.H;
        const int i = y * I.W + x;
        const bool carry = x0 == wx0 && y0 == wy0 + kRTY;
        top = carry ? ring_row(top + kRTY, kRWY) : 0;
        float a[24];
        if (valid) {
#pragma unroll
            for (int kb = 0, kk = 0; kb < 25; ++kb) {
                if (kb == 12) continue;
                a[kk++] = nt_at_off(I.L + (size_t)kb * I.pitch, 4u * (unsigned)i);
            }
        }
        float sv[kMaxRhs][kRNewPer];
#pragma unroll
        for (int j = 0; j < kMaxRhs; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const float* xs = sys[I.sys0 + j].x;
#pragma unroll
            for (int c = 0; c < kRNewPer; ++c) {
                const int e = (int)threadIdx.x + c * kBlock;
                const int wy = kKeep + e / kRWX, wx = e % kRWX;
                const long long g = (long long)(y0 - 2 + wy) * I.W + (x0 - 2 + wx);
                sv[j][c] = (e < kRNew && g >= 0 && g < I.N) ? at_off(xs, 4u * (unsigned)g) : 0.f;
            }
        }
        __syncthreads();  // the previous tile is done with win
        if (!carry) {
            for (int j = 0; j < kMaxRhs; ++j) {
                if (!((mask >> j) & 1u)) continue;
                const float* xs = sys[I.sys0 + j].x;
                for (int e = (int)threadIdx.x; e < kKeep * kRWX; e += kBlock) {
                    const int wy = e / kRWX, wx = e - wy * kRWX;
                    const long long g = (long long)(y0 - 2 + wy) * I.W + (x0 - 2 + wx);
                    win[j][ring_row(top + wy, kRWY) * kRWX + wx] =
                        (g >= 0 && g < I.N) ? at_off(xs, 4u * (unsigned)g) : 0.f;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kMaxRhs; ++j) {
            if (!((mask >> j) & 1u)) continue;
#pragma unroll
            for (int c = 0; c < kRNewPer; ++c) {
                const int e = (int)threadIdx.x + c * kBlock;
                if (e < kRNew) {
                    const int rr = e / kRWX;
                    win[j][ring_row(top + kKeep + rr, kRWY) * kRWX + (e - rr * kRWX)] = sv[j][c];
                }
            }
        }
        __syncthreads();
        wx0 = x0;
        wy0 = y0;
        if (!valid) continue;
        int rb[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) rb[s] = ring_row(top + row + s, kRWY) * kRWX + lane;
#pragma unroll
        for (int j = 0; j < kMaxRhs; ++j) {
            if (!((mask >> j) & 1u)) continue;
            const BSys& S = sys[I.sys0 + j];
            float* rcur = S.r[par[j]];
            const float* __restrict__ w = win[j];
            double s = 0.0;
#pragma unroll
            for (int kb = 0, kk = 0; kb < 25; ++kb) {
                if (kb == 12) continue;
                s = fma((double)a[kk], (double)w[rb[kb / 5] + (kb % 5)], s);
                ++kk;
            }
            const unsigned o = 4u * (unsigned)i;
            const float bi = at_off(S.b, o), di = at_off(S.d, o);
            const double tq = fma((double)di, (double)w[rb[2] + 2], s) - (double)bi;
            const float out = (float)tq;
            const float zo = inv_diag(di) * out;
            acc[j][0] += (double)out * (double)zo;
            acc[j][1] += tq * tq;
            if ((dmask >> j) & 1u) {
                const double dd = (double)out - (double)at_off(rcur, o);
                acc[j][2] += dd * dd;
            }
            acc[j][3] += (double)bi * (double)bi;
            if (op != 0) store_off(rcur, o, out);
        }
    }
    if (cur >= 0) flush_slot<kMaxRhs, 4>(acc, slot3 + slot_index(r, P3.K, cur, P3.wimg[r], 4 * kMaxRhs, 0));
}

// Finish a true-residual pass: one workgroup per image folds its slots and
// updates the selected systems' scalars; a replaced residual makes the
// system's next fused launch a first step.
__global__ __launch_bounds__(kBlock) void kb_trueres_fin(const BImg* __restrict__ imgs, Scalars* sc, BPart P3,
                                                         const double* __restrict__ slot3, int op, double tol) {
    const int im = blockIdx.x;
    const BImg I = imgs[im];
    for (int j = 0; j < I.nsys; ++j) {
        Scalars* s = &sc[I.sys0 + j];
        if (!tr_selected(s, op)) continue;
        double t[4];
        img_fold<4>(slot3, P3, im, I.wf3, I.wl3, 4 * kMaxRhs, 4 * j, t);
        if (threadIdx.x != 0) continue;
        const double rz = t[0], rr = t[1], d2 = t[2], bb = t[3];
        s->g0 = rr;
        if (op == 0) continue;
        if (op == 2) {
            s->bb = bb;
            s->tol2 = tol > 0 ? tol * tol * (bb > 0 ? bb : 1.0) : -1.0;
        }
        replace_residual(s, rz, rr, op == 2, d2, rr);
        s->fresh = s->done == kRun ? 1 : 0;
    }
}

// per system: d = L_ii + lam, b = lam*tgt, x = x0 or 0
__global__ __launch_bounds__(kBlock) void kb_set_rhs(const float* __restrict__ Ldiag, int N,
                                                     const float* __restrict__ lam, const float* __restrict__ tgt,
                                                     const float* __restrict__ x0, float* __restrict__ d,
                                                     float* __restrict__ b, float* __restrict__ x) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        const float l = lam ? lam[i] : 0.f;
        float di = Ldiag[i];
        float bi = 0.f;
        if (l != 0.f) {
            di += l;
            bi = l * (tgt ? tgt[i] : 0.f);
        }
        d[i] = di;
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
                const int kk = 4 * q + h, kb = kk < 12 ? kk : kk + 1;
                v[h] = L[(size_t)kb * pitch + i];
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

// ---------------------------------------------------------------------------
// launch tables: kb_fused<FMT, NR, TX> for the batch's format, right-hand
// sides rounded up to a power of two, and tile width

typedef void (*FusedFn)(const BImg*, Scalars*, BPart, double*, size_t, int, int, int);
constexpr int kTileW[3] = {60, 124, 252};  // tile widths (windows 64, 128, 256 wide)
constexpr int kNumTileW = 3;

template <int FMT, int NR>
static FusedFn fused_fn_nr(int tw) {
    if (tw == 60) return kb_fused<FMT, NR, 64>;
    if (tw == 124) return kb_fused<FMT, NR, 128>;
    if constexpr (NR <= 4) {  // 8 systems of 256-wide windows exceed 64 KB of LDS
        if (tw == 252) return kb_fused<FMT, NR, 256>;
    }
    return nullptr;
}
template <int FMT>
static FusedFn fused_fn_fmt(int nr, int tw) {
    switch (nr) {
        case 1: return fused_fn_nr<FMT, 1>(tw);
        case 2: return fused_fn_nr<FMT, 2>(tw);
        case 4: return fused_fn_nr<FMT, 4>(tw);
        case 8: return fused_fn_nr<FMT, 8>(tw);
    }
    return nullptr;
}
static FusedFn fused_fn(int precision, int nr, int tw) {
    if (precision == GM_PREC_FP16) return fused_fn_fmt<kF16>(nr, tw);
    if (precision == GM_PREC_BF16) return fused_fn_fmt<kBF16>(nr, tw);
    return fused_fn_fmt<kF32G>(nr, tw);
}

typedef void (*UpdFn)(const BImg*, const BSys*, Scalars*, BPart, BPart, const double*, size_t, int, int);
typedef void (*FoldFn)(const BImg*, Scalars*, BPart, const double*, size_t, int, int);
static UpdFn upd_fn(int nr) {
    switch (nr) {
        case 1: return kb_fupdate<1>;
        case 2: return kb_fupdate<2>;
        case 4: return kb_fupdate<4>;
        default: return kb_fupdate<8>;
    }
}
static FoldFn fold_fn(int nr) {
    switch (nr) {
        case 1: return kb_ffold<1>;
        case 2: return kb_ffold<2>;
        case 4: return kb_ffold<4>;
        default: return kb_ffold<8>;
    }
}

}  // namespace gm

using namespace gm;

// Host side of a partition: the plan plus its device copies.
struct HostPart {
    bplan::Part plan;
    int* dimg = nullptr;
    int* dwimg = nullptr;
    BPart dev() const { return BPart{plan.T, plan.G, plan.K, dimg, dwimg}; }
};

struct gm_batch_s {
    int nimg = 0, nrhs = 0, nr = 1, nsys = 0, precision = 0;
    std::vector<BImg> himg;
    std::vector<BSys> hsys;
    BImg* dimg = nullptr;
    BSys* dsys = nullptr;
    Scalars* dsc = nullptr;
    Scalars* hsc = nullptr;   // pinned, nsys
    HostPart pf, pq, pc;      // fused tiles, true-residual tiles, update-only chunks
    double* dslot = nullptr;  // fused sums (two parity halves of fhalf), then the true-residual sums
    size_t fhalf = 0, slot3_n = 0;
    float* Lmem = nullptr;
    unsigned short* L16mem = nullptr;
    float* Lgmem = nullptr;  // grouped FP32 bands (FP32 batches)
    float* vmem = nullptr;
    hipStream_t cap = nullptr;
    // a chunk of fused launches: [0] ends with the fold that publishes the
    // state, [1] (before a residual replacement) with an update-only launch
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int graph_iters = 0, graph_verify = -1;
    hipEvent_t ev = nullptr;
    int per_cu = 32;   // workgroups per CU of the fused launch (gm_batch_set_grid, gm_batch_autotune)
    int tile_w = 124;  // fused tile width (gm_batch_set_tiles, gm_batch_autotune)
    double tuned_us = 0;
};

static void free_part(HostPart& p) {
    if (p.dimg) (void)hipFree(p.dimg);
    if (p.dwimg) (void)hipFree(p.dwimg);
    p.dimg = p.dwimg = nullptr;
}

static void drop_graphs(gm_batch* b) {
    for (int g = 0; g < 2; ++g) {
        if (b->gexec[g]) (void)hipGraphExecDestroy(b->gexec[g]);
        if (b->graph[g]) (void)hipGraphDestroy(b->graph[g]);
        b->gexec[g] = nullptr;
        b->graph[g] = nullptr;
    }
    b->graph_iters = 0;
}

static void batch_free(gm_batch* b) {
    if (!b) return;
    drop_graphs(b);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->ev) (void)hipEventDestroy(b->ev);
    free_part(b->pf);
    free_part(b->pq);
    free_part(b->pc);
    for (void* p : {(void*)b->dimg, (void*)b->dsys, (void*)b->dsc, (void*)b->dslot, (void*)b->Lmem,
                    (void*)b->L16mem, (void*)b->Lgmem, (void*)b->vmem})
        if (p) (void)hipFree(p);
    if (b->hsc) (void)hipHostFree(b->hsc);
    delete b;
}

// Workgroups of a batched launch: per_cu per CU (more than fit at once: a
// workgroup whose range lies in converged images ends at once and the
// hardware starts the next, which balances the uneven work), at most one per
// item, a multiple of the XCD count, at most kBatchMaxGrid (the batched
// kernels' slot arrays are sized by the grid they run on: 8192 workgroups, 32
// per CU, shorten the ranges the last images to converge are split into).
static int batch_grid(const DevInfo& d, long items, int per_cu) {
    long g = std::min((long)d.cus * per_cu, items);
    g = std::max(g, 1L);
    g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
    return (int)std::min<long>(g, kBatchMaxGrid);
}

static gm_status upload_part(HostPart& p) {
    free_part(p);
    GM_HIP(hipMalloc(&p.dimg, sizeof(int) * std::max(p.plan.T, 1)));
    GM_HIP(hipMalloc(&p.dwimg, sizeof(int) * p.plan.G));
    GM_HIP(hipMemcpy(p.dimg, p.plan.img.data(), sizeof(int) * p.plan.T, hipMemcpyHostToDevice));
    GM_HIP(hipMemcpy(p.dwimg, p.plan.wimg.data(), sizeof(int) * p.plan.G, hipMemcpyHostToDevice));
    return GM_OK;
}

// (Re)build the three work lists and partitions for the batch's tile width
// and grid, and upload them.
static gm_status batch_partition(gm_batch* b) {
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    std::vector<int> W(b->nimg), H(b->nimg);
    for (int i = 0; i < b->nimg; ++i) {
        W[i] = b->himg[i].W;
        H[i] = b->himg[i].H;
    }
    std::vector<bplan::ImgLists> li;
    bplan::build_lists(W, H, b->tile_w, kBlock / (b->tile_w + 4), kRTX, kRTY, kChunk, li, b->pf.plan, b->pq.plan,
                       b->pc.plan);
    for (int i = 0; i < b->nimg; ++i) {
        BImg& I = b->himg[i];
        I.t0 = li[i].t0;
        I.nty = li[i].nty;
        I.q0 = li[i].q0;
        I.nqy = li[i].nqy;
        I.c0 = li[i].c0;
    }
    bplan::split(b->pf.plan, batch_grid(d, (long)b->pf.plan.img.size(), b->per_cu), b->nimg);
    bplan::split(b->pq.plan, batch_grid(d, (long)b->pq.plan.img.size(), 16), b->nimg);
    bplan::split(b->pc.plan, batch_grid(d, (long)b->pc.plan.img.size(), 32), b->nimg);
    const int segf = bplan::seg_len(b->pf.plan.T, b->pf.plan.G), segq = bplan::seg_len(b->pq.plan.T, b->pq.plan.G);
    for (int i = 0; i < b->nimg; ++i) {
        BImg& I = b->himg[i];
        I.segf = segf;
        I.segq = segq;
        I.wf1 = b->pf.plan.wf[i];
        I.wl1 = b->pf.plan.wl[i];
        I.wf3 = b->pq.plan.wf[i];
        I.wl3 = b->pq.plan.wl[i];
        I.wf2 = b->pc.plan.wf[i];
    }
    for (HostPart* p : {&b->pf, &b->pq, &b->pc})
        if ((st = upload_part(*p)) != GM_OK) return st;
    b->fhalf = (size_t)b->pf.plan.G * b->pf.plan.K * kNQ * b->nr;
    b->slot3_n = (size_t)b->pq.plan.G * b->pq.plan.K * 4 * kMaxRhs;
    if (b->dslot) (void)hipFree(b->dslot);
    b->dslot = nullptr;
    const size_t n = 2 * b->fhalf + b->slot3_n;
    GM_HIP(hipMalloc(&b->dslot, sizeof(double) * n));
    GM_HIP(hipMemset(b->dslot, 0, sizeof(double) * n));
    GM_HIP(hipMemcpy(b->dimg, b->himg.data(), sizeof(BImg) * b->nimg, hipMemcpyHostToDevice));
    drop_graphs(b);  // a captured graph bakes in the partition
    return GM_OK;
}

extern "C" {

gm_status gm_batch_create(gm_batch** out, int nimages, const int* widths, const int* heights, int nrhs, int precision) {
    if (!out || nimages <= 0 || !widths || !heights || nrhs <= 0 || nrhs > kMaxRhs)
        return fail(GM_ERR_INVALID, "gm_batch_create: bad arguments (nrhs must be 1..%d)", kMaxRhs);
    if (precision != GM_PREC_FP32 && precision != GM_PREC_FP16 && precision != GM_PREC_BF16)
        return fail(GM_ERR_UNSUPPORTED, "gm_batch_create: batched forms are FP32, FP16 and BF16");
    for (int i = 0; i < nimages; ++i) {
        const int W = widths[i], H = heights[i];
        // below 2^25 pixels: a system's block of 9 vectors spans less than
        // 2^31 bytes (one range-checked buffer descriptor per system in the
        // fused iteration), and the grouped band loads' 32-bit byte offsets
        // (16 bytes per pixel) fit
        if (W < 1 || H < 1 || (long long)W * H >= (1LL << 25))
            return fail(GM_ERR_INVALID, "gm_batch_create: image %d is %dx%d (1 to 2^25-1 pixels)", i, W, H);
    }
    gm_batch* b = new gm_batch();
    b->nimg = nimages;
    b->nrhs = nrhs;
    b->nr = nrhs <= 1 ? 1 : nrhs <= 2 ? 2 : nrhs <= 4 ? 4 : 8;
    b->nsys = nimages * nrhs;
    b->precision = precision;
    size_t Lfloats = 0, L16n = 0, Lgn = 0, vfloats = 0;
    std::vector<size_t> loff(nimages), l16off(nimages), lgoff(nimages), voff(nimages);
    for (int i = 0; i < nimages; ++i) {
        const int W = widths[i], H = heights[i];
        const int N = W * H;
        const size_t pitch = (size_t)(N + 63) / 64 * 64;
        BImg I{};
        I.W = W;
        I.H = H;
        I.N = N;
        I.sys0 = i * nrhs;
        I.nsys = nrhs;
        I.pitch = pitch;
        loff[i] = Lfloats;
        Lfloats += 25 * pitch;
        l16off[i] = L16n;
        if (precision != GM_PREC_FP32) L16n += 24 * pitch;
        lgoff[i] = Lgn;
        if (precision == GM_PREC_FP32) Lgn += 24 * pitch;
        voff[i] = vfloats;
        vfloats += (size_t)nrhs * kSysVecs * pitch;
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
        I.V = b->vmem + voff[i];
        for (int j = 0; j < nrhs; ++j) {
            float* base = I.V + (size_t)j * kSysVecs * I.pitch;
            BSys S;
            S.x = base;
            S.r[0] = base + 1 * I.pitch;
            S.r[1] = base + 2 * I.pitch;
            S.p[0] = base + 3 * I.pitch;
            S.p[1] = base + 4 * I.pitch;
            S.Ap[0] = base + 5 * I.pitch;
            S.Ap[1] = base + 6 * I.pitch;
            S.d = base + 7 * I.pitch;
            S.b = base + 8 * I.pitch;
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
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    const dim3 grid(elementwise_grid(d, I.N, 1));
    if (b->precision == GM_PREC_FP32)
        hipLaunchKernelGGL(kb_group32, grid, dim3(kBlock), 0, (hipStream_t)s, I.L, I.pitch, I.N,
                           const_cast<float*>(I.Lg));
    else if (b->precision == GM_PREC_BF16)
        hipLaunchKernelGGL(kb_to16<kBF16>, grid, dim3(kBlock), 0, (hipStream_t)s, I.L, I.pitch, I.N,
                           const_cast<unsigned short*>(I.L16), I.pitch);
    else
        hipLaunchKernelGGL(kb_to16<kF16>, grid, dim3(kBlock), 0, (hipStream_t)s, I.L, I.pitch, I.N,
                           const_cast<unsigned short*>(I.L16), I.pitch);
    GM_LAUNCH_CHECK();
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
                       I.L + 12 * I.pitch, I.N, lam, tgt, x0, S.d, S.b, S.x);
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

static double* fslots(gm_batch* b) { return b->dslot; }
static double* slot3(gm_batch* b) { return b->dslot + 2 * b->fhalf; }

static gm_status launch_fused(gm_batch* b, int k, int tune, int verify, hipStream_t hs) {
    FusedFn f = fused_fn(b->precision, b->nr, b->tile_w);
    if (!f) return fail(GM_ERR_UNSUPPORTED, "gm_batch: no %d-wide tiles for %d right-hand sides", b->tile_w, b->nr);
    hipLaunchKernelGGL(f, dim3(b->pf.plan.G), dim3(kBlock), 0, hs, b->dimg, b->dsc, b->pf.dev(), fslots(b), b->fhalf,
                       k, tune, verify);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status launch_fupdate(gm_batch* b, int k, int verify, hipStream_t hs) {
    hipLaunchKernelGGL(upd_fn(b->nr), dim3(b->pc.plan.G), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, b->pf.dev(),
                       b->pc.dev(), fslots(b), b->fhalf, k, verify);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status launch_ffold(gm_batch* b, int par, int verify, hipStream_t hs) {
    hipLaunchKernelGGL(fold_fn(b->nr), dim3(b->nimg), dim3(kBlock), 0, hs, b->dimg, b->dsc, b->pf.dev(), fslots(b),
                       b->fhalf, par, verify);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status batch_trueres(gm_batch* b, int op, double tol, hipStream_t hs) {
    hipLaunchKernelGGL(kb_trueres, dim3(b->pq.plan.G), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, b->pq.dev(), op,
                       slot3(b));
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_trueres_fin, dim3(b->nimg), dim3(kBlock), 0, hs, b->dimg, b->dsc, b->pq.dev(), slot3(b), op,
                       tol);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

// The chunks of `iters` (even, >= 2) fused launches, captured on the batch's
// own stream (capture needs a non-null stream) and replayed on the caller's.
static gm_status batch_graphs(gm_batch* b, int iters, int verify) {
    if (b->gexec[0] && b->graph_iters == iters && b->graph_verify == verify) return GM_OK;
    drop_graphs(b);
    for (int g = 0; g < 2; ++g) {
        GM_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        gm_status st = GM_OK;
        const int nfull = g == 0 ? iters : iters - 1;
        for (int k = 0; k < nfull && st == GM_OK; ++k) st = launch_fused(b, k, 0, verify, b->cap);
        if (st == GM_OK)
            st = g == 0 ? launch_ffold(b, iters & 1, verify, b->cap) : launch_fupdate(b, iters - 1, verify, b->cap);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap, &gr);
        if (st != GM_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return st;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        b->graph[g] = gr;
        GM_HIP(hipGraphInstantiate(&b->gexec[g], gr, nullptr, nullptr, 0));
    }
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
    int chunk = o.check_every > 0 ? o.check_every : 10;
    chunk = std::max(2, chunk + (chunk & 1));  // even, >= 2: the launch parities of one chunk continue in the next
    const bool mixed = b->precision != GM_PREC_FP32;
    int refine = o.refine_every > 0 ? o.refine_every : (o.refine_every == 0 && mixed ? 50 : 0);
    if (refine > 0 && refine % chunk) refine = (refine / chunk + 1) * chunk;  // replacement at chunk boundaries

    hipLaunchKernelGGL(kb_init, dim3((b->nsys + 63) / 64), dim3(64), 0, hs, b->dsc, b->nsys, o.max_iter);
    GM_LAUNCH_CHECK();
    if ((st = batch_trueres(b, 2, o.tol, hs)) != GM_OK) return st;  // r = A x - b; ||b||^2 -> tol2
    if ((st = batch_graphs(b, chunk, verify)) != GM_OK) return st;

    long enq = 0;
    bool finished = false;
    while (!finished) {
        enq += chunk;
        const bool rep = refine > 0 && enq % refine == 0;
        GM_HIP(hipGraphLaunch(b->gexec[rep ? 1 : 0], hs));
        if (rep && (st = batch_trueres(b, 1, o.tol, hs)) != GM_OK) return st;
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
        if (verifying > 0 && (st = batch_trueres(b, 3, o.tol, hs)) != GM_OK) return st;
    }
    // the FP64 true residual of every system (g0) and a host copy
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

gm_status gm_batch_set_tiles(gm_batch* b, int tile_width) {
    if (!b) return fail(GM_ERR_INVALID, "gm_batch_set_tiles: null batch");
    if (!fused_fn(b->precision, b->nr, tile_width))
        return fail(GM_ERR_INVALID, "gm_batch_set_tiles: width %d (60, 124 or 252; 252 needs <= 4 right-hand sides)",
                    tile_width);
    b->tile_w = tile_width;
    return batch_partition(b);
}

// The tuner's candidates: every tile width, 4 to 32 workgroups per CU.  Each
// is timed by `reps` tuning launches at full load (every system steps with
// fixed alpha and beta: all the loads and stores of a real launch), or, with
// solve options, by `reps` whole solves -- which also weighs the tail of a
// solve, where only the slowest images still run and more, shorter ranges
// keep more CUs busy.  The initial guesses are saved before and restored
// after; the next gm_batch_solve rebuilds the rest of the state.
gm_status gm_batch_autotune(gm_batch* b, const gm_solve_opts* opts, int reps, double* best_us, gm_stream_t s) {
    if (!b || reps < 1) return fail(GM_ERR_INVALID, "gm_batch_autotune: arguments");
    hipStream_t hs = (hipStream_t)s;
    size_t nx = 0;
    for (const BImg& I : b->himg) nx += (size_t)I.nsys * I.N;
    float* save = nullptr;
    GM_HIP(hipMalloc(&save, sizeof(float) * std::max<size_t>(nx, 1)));
    auto copy_x = [&](bool to_save) -> gm_status {
        size_t o = 0;
        for (const BImg& I : b->himg)
            for (int j = 0; j < I.nsys; ++j) {
                float* x = b->hsys[I.sys0 + j].x;
                GM_HIP(hipMemcpyAsync(to_save ? save + o : x, to_save ? x : save + o, sizeof(float) * I.N,
                                      hipMemcpyDeviceToDevice, hs));
                o += I.N;
            }
        return GM_OK;
    };
    gm_status st = copy_x(true);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st == GM_OK && hipEventCreate(&e0) != hipSuccess) st = fail(GM_ERR_HIP, "hipEventCreate");
    if (st == GM_OK && hipEventCreate(&e1) != hipSuccess) st = fail(GM_ERR_HIP, "hipEventCreate");
    const int keep_w = b->tile_w, keep_pc = b->per_cu;
    int bw = keep_w, bpc = keep_pc;
    double best = 1e300;
    const int pcs[] = {4, 8, 16, 32};
    for (int wi = 0; wi < kNumTileW && st == GM_OK; ++wi) {
        if (!fused_fn(b->precision, b->nr, kTileW[wi])) continue;
        for (int pc : pcs) {
            b->tile_w = kTileW[wi];
            b->per_cu = pc;
            if ((st = batch_partition(b)) != GM_OK) break;
            double us = 0;
            if (opts) {
                for (int r = 0; r <= reps && st == GM_OK; ++r) {  // r = 0: warm-up (graph capture)
                    if ((st = copy_x(false)) != GM_OK) break;
                    if (hipEventRecord(e0, hs) != hipSuccess) st = fail(GM_ERR_HIP, "hipEventRecord");
                    gm_status sv = st == GM_OK ? gm_batch_solve(b, opts, s) : st;
                    if (sv != GM_OK && sv != GM_ERR_NOT_CONVERGED) st = sv;
                    if (st == GM_OK && (hipEventRecord(e1, hs) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                        st = fail(GM_ERR_HIP, "hipEventSynchronize");
                    float ms = 0.f;
                    if (st == GM_OK && r > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) us += 1e3 * ms;
                }
            } else {
                if ((st = launch_fused(b, 0, 1, 0, hs)) != GM_OK) break;  // warm-up
                if (hipEventRecord(e0, hs) != hipSuccess) st = fail(GM_ERR_HIP, "hipEventRecord");
                for (int r = 0; r < reps && st == GM_OK; ++r) st = launch_fused(b, 0, 1, 0, hs);
                if (st == GM_OK && (hipEventRecord(e1, hs) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                    st = fail(GM_ERR_HIP, "hipEventSynchronize");
                float ms = 0.f;
                if (st == GM_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) us = 1e3 * ms;
            }
            if (st != GM_OK) break;
            us /= reps;
            if (us < best) {
                best = us;
                bw = kTileW[wi];
                bpc = pc;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    b->tile_w = st == GM_OK ? bw : keep_w;
    b->per_cu = st == GM_OK ? bpc : keep_pc;
    gm_status st2 = batch_partition(b);
    if (st == GM_OK) st = st2;
    if (st == GM_OK) st = copy_x(false);
    if (hipStreamSynchronize(hs) != hipSuccess && st == GM_OK) st = fail(GM_ERR_HIP, "hipStreamSynchronize");
    (void)hipFree(save);
    if (st == GM_OK) {
        b->tuned_us = best;
        if (best_us) *best_us = best;
    }
    return st;
}

gm_status gm_batch_get_config(gm_batch* b, int* tile_width, int* per_cu, double* tuned_us) {
    if (!b) return fail(GM_ERR_INVALID, "gm_batch_get_config: null batch");
    if (tile_width) *tile_width = b->tile_w;
    if (per_cu) *per_cu = b->per_cu;
    if (tuned_us) *tuned_us = b->tuned_us;
    return GM_OK;
}

int gm_batch_tiles(gm_batch* b) { return b ? b->pf.plan.T : 0; }

}  // extern "C"
