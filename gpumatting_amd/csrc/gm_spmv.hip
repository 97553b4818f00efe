// gm_spmv.hip -- banded matrix-vector products (BandedMatrix.cu:61-207 roles).
//
// Two kernels compute y = A x (or A x +- y):
//   * k_spmv<Geometry>: the Levin 5x5-band structure (offsets dy*W+dx).  Each
//     thread first issues the 16-byte non-temporal loads of its R rows' 25
//     diagonals, the workgroup then stages its tile's x window in LDS with
//     coalesced loads (a 2-D pixel halo, or 5 row segments for the 1-D
//     geometry), and the products read x from LDS.  Persistent, XCD-aware
//     grid; tile shape and grid size from the autotuner's table.
//     HBM-bound: N*(25*s_A + 8) algorithmic bytes.
//   * k_bm_generic: any band set (<= 64 bands), one row per thread, x read
//     directly at i + off_k (caller's zero padding, as in the reference).
#include <map>
#include <utility>

#include <string>

#include "gm_internal.h"

namespace gm {

// MODE 0: y = A x; 1: y = A x + z; 2: y = A x - z
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_bm_generic(float* __restrict__ out, const float* __restrict__ a,
                                                       size_t pitch, BandArgs ba, int rows,
                                                       const float* __restrict__ x, const float* __restrict__ z) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < rows; i += gridDim.x * kBlock) {
        float s = 0.f;
        for (int k = 0; k < ba.nb; ++k) s = fmaf(a[(size_t)i + (size_t)k * pitch], x[(long long)i + ba.off[k]], s);
        if (MODE == 1) s += z[i];
        if (MODE == 2) s -= z[i];
        out[i] = s;
    }
}

template <class G, int FMT, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv(float* __restrict__ out, const void* __restrict__ A, size_t pitch,
                                                 int W, int H, int N, const float* __restrict__ x,
                                                 const float* __restrict__ z) {
    constexpr int R = G::R;
    __shared__ __attribute__((aligned(16))) float lds[G::LDS_FLOATS];
    const TileRange tr = xcd_tiles(G::ntiles(W, H, N));
    if constexpr (G::kIs2D) {
        // Software pipeline: tile t + step's bands and x window are loaded
        // into registers while tile t is computed.
        constexpr int NE = G::LDS_FLOATS;  // x window cells
        constexpr int XE = (NE + kBlock - 1) / kBlock;
        Bands<R, FMT> nb;
        float nx[XE];
        auto issue = [&](int t) {
            G g;
            g.at(t, W, H, N);
            if (g.valid) load_bands<R, FMT, kDiagStored>(A, pitch, g.ib, nb);
#pragma unroll
            for (int q = 0; q < XE; ++q) {
                const int e = (int)threadIdx.x + q * kBlock;
                const int ly = e / G::STRIDE, lx = e - ly * G::STRIDE;
                // matrix-row index of the window cell (as Geom2D::stage)
                const long long gi = (long long)(g.y0 - 2 + ly) * W + (g.x0 - 2 + lx);
                nx[q] = (e < NE && gi >= 0 && gi < N) ? x[gi] : 0.f;
            }
        };
        if (tr.begin < tr.end) issue(tr.begin);
        for (int t = tr.begin; t < tr.end; t += tr.step) {
            G g;
            g.at(t, W, H, N);
            __syncthreads();  // the previous tile is done with lds
#pragma unroll
            for (int q = 0; q < XE; ++q) {
                const int e = (int)threadIdx.x + q * kBlock;
                if (e < NE) lds[e] = nx[q];
            }
            const Bands<R, FMT> av = nb;
            __syncthreads();
            if (t + tr.step < tr.end) issue(t + tr.step);
            if (g.valid) {
                float y[R];
                apply_bands<R, FMT, kDiagStored, G::STRIDE>(av, g.xs(lds), y);
                if (MODE != 0) {
                    float zz[R];
                    load_rows<R>(z, g.ib, N, zz);
#pragma unroll
                    for (int j = 0; j < R; ++j) y[j] = MODE == 1 ? y[j] + zz[j] : y[j] - zz[j];
                }
                store_rows<R>(out, g.ib, N, y);
            }
        }
    } else {
        for (int t = tr.begin; t < tr.end; t += tr.step) {
            G g;
            g.at(t, W, H, N);
            Bands<R, FMT> av;
            if (g.valid) load_bands<R, FMT, kDiagStored>(A, pitch, g.ib, av);
            __syncthreads();
            g.stage(lds, W, H, N, [&](int i) { return x[i]; });
            __syncthreads();
            if (g.valid) {
                float y[R];
                apply_bands<R, FMT, kDiagStored, G::STRIDE>(av, g.xs(lds), y);
                if (MODE != 0) {
                    float zz[R];
                    load_rows<R>(z, g.ib, N, zz);
#pragma unroll
                    for (int j = 0; j < R; ++j) y[j] = MODE == 1 ? y[j] + zz[j] : y[j] - zz[j];
                }
                store_rows<R>(out, g.ib, N, y);
            }
        }
    }
}

// FP32 banded SpMV without LDS: each thread reads its own x window straight
// from global memory (neighbouring threads share it in L1 / L2), so there is
// no staging and no barrier, and fewer registers allow more waves per CU.
// 2-D tiles of TX x TY pixels, R consecutive pixels per thread (W % R == 0,
// x 4R-byte aligned).  Tiles whose window reaches the first 3 or last 3
// image rows guard every x read; the others read x at matrix-row indices
// that stay inside [0, N) (cells beside the image alias the neighbouring row
// exactly like the row-wise product; their band entries are 0).
template <int R, int TY, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_direct(float* __restrict__ out, const float* __restrict__ A,
                                                        size_t pitch, int W, int H, int N,
                                                        const float* __restrict__ x, const float* __restrict__ z) {
    static_assert(R == 2 || R == 4, "rows per thread");
    using G = Geom2D<R, TY>;
    constexpr int XL = R == 4 ? 12 : 6;  // x floats read per window row, from column x - (R == 4 ? 4 : 2)
    constexpr int CO = R == 4 ? 2 : 0;   // index of column x - 2 among them
    typedef float vecR __attribute__((ext_vector_type(R)));
    const TileRange tr = xcd_tiles(G::ntiles(W, H, N));
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        G g;
        g.at(t, W, H, N);
        if (!g.valid) continue;
        Bands<R, kF32> av;
        load_bands<R, kF32, kDiagStored>(A, pitch, g.ib, av);
        const bool interior = g.y0 >= 3 && g.y0 + TY + 3 <= H;
        float xw[5][XL];
#pragma unroll
        for (int sr = 0; sr < 5; ++sr) {
            const long long base = (long long)g.ib + (long long)(sr - 2) * W - (R == 4 ? 4 : 2);
            if (interior) {
#pragma unroll
                for (int q = 0; q < XL / R; ++q) {
                    const vecR v = *reinterpret_cast<const vecR*>(x + base + q * R);
#pragma unroll
                    for (int c = 0; c < R; ++c) xw[sr][q * R + c] = v[c];
                }
            } else {
#pragma unroll
                for (int c = 0; c < XL; ++c) {
                    const long long gi = base + c;
                    xw[sr][c] = (gi >= 0 && gi < N) ? x[gi] : 0.f;
                }
            }
        }
        float y[R];
#pragma unroll
        for (int j = 0; j < R; ++j) y[j] = 0.f;
#pragma unroll
        for (int k = 0; k < 25; ++k) {
            const int sr = k / 5, dx = k % 5;
#pragma unroll
            for (int j = 0; j < R; ++j) y[j] = fmaf(av.v[k].get(j), xw[sr][CO + dx + j], y[j]);
        }
        if (MODE != 0) {
            float zz[R];
            load_rows<R>(z, g.ib, N, zz);
#pragma unroll
            for (int j = 0; j < R; ++j) y[j] = MODE == 1 ? y[j] + zz[j] : y[j] - zz[j];
        }
        store_rows<R>(out, g.ib, N, y);
    }
}

// FP32 banded SpMV on long row runs: tiles of TX = 256 G pixels by 4 image
// rows, one wave per tile row, and each lane G runs of 4 pixels, 1 KB apart,
// so a wave reads G KB contiguous of a band plane where the 2-D tiles read
// 1 KB.  The bands are read in five passes of five (window row dy, then dx:
// band order k = 5 dy + dx), each pass's loads issued before its products, so
// a wave has 5 G loads in flight against 25 of 1 KB each.  Streaming 25 planes
// with 4 KB runs per wave, five planes at a time, measured 6.2 TB/s against
// 5.7 TB/s for 1 KB runs of all 25 at once.  Every output is the FMA chain of
// apply_bands (k = 0..24 from 0), so the result is bit-identical to the other
// forms.  The x window, (4 + 4) x (TX + 4) cells addressed by matrix-row
// index, is staged in LDS with window column c at LDS column c, so each run's
// 8 window cells are two aligned 16-byte reads.  Needs W % 4 == 0, 16-byte
// aligned out, x and z, and 25 band planes below 2^31 bytes.
template <int G, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_seg(float* __restrict__ out, const float* __restrict__ A,
                                                     size_t pitch, int W, int H, int N, const float* __restrict__ x,
                                                     const float* __restrict__ z) {
    constexpr int TX = 256 * G, TY = 4, ST = TX + 4, C4 = TX / 4;
    constexpr int IE = (TY + 4) * C4 / kBlock;  // interior float4 cells staged per thread
    static_assert((TY + 4) * C4 % kBlock == 0, "whole staging rounds");
    __shared__ __attribute__((aligned(16))) float lds[(TY + 4) * ST];
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int ncol = (W + TX - 1) / TX, nrow = (H + TY - 1) / TY;
    const TileRange tr = xcd_tiles(ncol * nrow);
    const __amdgpu_buffer_rsrc_t bx = buf_rsrc(x, N), bz = buf_rsrc(z, MODE ? N : 0), bo = buf_rsrc(out, N);
    const __amdgpu_buffer_rsrc_t ba =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(A), 0, (int)(25 * 4 * pitch), 0x00020000);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int x0 = (t % ncol) * TX, y0 = (t / ncol) * TY, y = y0 + w;
        unsigned off[G];  // byte offset of each run in a band plane (out of range: invalid)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int xc = x0 + 256 * g + 4 * lane;
            off[g] = (xc < W && y < H) ? 4u * (unsigned)(y * W + xc) : 0x80000000u;
        }
        f32x4 a[5][G];
        auto issue = [&](int dy) {
#pragma unroll
            for (int dx = 0; dx < 5; ++dx)
#pragma unroll
                for (int g = 0; g < G; ++g)
                    a[dx][g] = __builtin_bit_cast(
                        f32x4, __builtin_amdgcn_raw_buffer_load_b128(ba, (int)off[g], (int)(4 * (5 * dy + dx) * pitch),
                                                                     kNonTemporal));
        };
        __syncthreads();  // the previous tile is done with lds
        f32x4 xi[IE];
#pragma unroll
        for (int q = 0; q < IE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock, ly = e / C4, c4 = e - ly * C4;
            xi[q] = buf_ld4(bx, (long long)(y0 - 2 + ly) * W + x0 + 4 * c4);
        }
        float xh = 0.f;
        const int hly = (int)threadIdx.x >> 2, hc = (int)threadIdx.x & 3, hlx = hc < 2 ? hc : TX + hc;
        if (threadIdx.x < 4 * (TY + 4)) xh = buf_ld(bx, (long long)(y0 - 2 + hly) * W + x0 - 2 + hlx);
        issue(0);
#pragma unroll
        for (int q = 0; q < IE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock, ly = e / C4, c4 = e - ly * C4;
            float* d = lds + ly * ST + 2 + 4 * c4;
            *reinterpret_cast<f32x2*>(d) = f32x2{xi[q].x, xi[q].y};
            *reinterpret_cast<f32x2*>(d + 2) = f32x2{xi[q].z, xi[q].w};
        }
        if (threadIdx.x < 4 * (TY + 4)) lds[hly * ST + hlx] = xh;
        __syncthreads();
        f32x4 acc[G];
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int dy = 0; dy < 5; ++dy) {
            if (dy > 0) issue(dy);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float* s = lds + (w + dy) * ST + 256 * g + 4 * lane;
                const f32x4 lo = *reinterpret_cast<const f32x4*>(s), hi = *reinterpret_cast<const f32x4*>(s + 4);
                const float xw[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
                for (int dx = 0; dx < 5; ++dx)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[g][j] = fmaf(a[dx][g][j], xw[j + dx], acc[g][j]);
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            f32x4 v = acc[g];
            if constexpr (MODE != 0) {
                const f32x4 zz =
                    __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(bz, (int)off[g], 0, 0));
                v = MODE == 1 ? v + zz : v - zz;
            }
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), bo, (int)off[g], 0, 0);
        }
    }
}

template <int G>
static void spmv_seg_t(float* out, const float* A, size_t pitch, int W, int N, const float* x, const float* z,
                       int mode, int grid, hipStream_t hs) {
    const int H = N / W;
    if (mode == 0)
        hipLaunchKernelGGL((k_spmv_seg<G, 0>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
    else if (mode == 1)
        hipLaunchKernelGGL((k_spmv_seg<G, 1>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
    else
        hipLaunchKernelGGL((k_spmv_seg<G, 2>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
}

template <int R, int TY>
static void spmv_direct_t(float* out, const float* A, size_t pitch, int W, int N, const float* x, const float* z,
                          int mode, int grid, hipStream_t hs) {
    const int H = N / W;
    if (mode == 0)
        hipLaunchKernelGGL((k_spmv_direct<R, TY, 0>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
    else if (mode == 1)
        hipLaunchKernelGGL((k_spmv_direct<R, TY, 1>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
    else
        hipLaunchKernelGGL((k_spmv_direct<R, TY, 2>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
}

template <class G, int FMT>
static void spmv_t(float* out, const void* A, size_t pitch, int W, int N, const float* x, const float* z, int mode,
                   int grid, hipStream_t hs) {
    const int H = G::kIs2D ? N / W : 0;
    if (mode == 0)
        hipLaunchKernelGGL((k_spmv<G, FMT, 0>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
    else if (mode == 1)
        hipLaunchKernelGGL((k_spmv<G, FMT, 1>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
    else
        hipLaunchKernelGGL((k_spmv<G, FMT, 2>), dim3(grid), dim3(kBlock), 0, hs, out, A, pitch, W, H, N, x, z);
}

template <int FMT>
static void spmv_shape_t(float* out, const void* A, size_t pitch, int W, int N, const float* x, const float* z,
                         int mode, int shape, int grid, hipStream_t hs) {
    switch (shape) {
        case kS1R2: spmv_t<Geom1D<2>, FMT>(out, A, pitch, W, N, x, z, mode, grid, hs); break;
        case kS1R4: spmv_t<Geom1D<4>, FMT>(out, A, pitch, W, N, x, z, mode, grid, hs); break;
        case kS1R8: spmv_t<Geom1D<8>, FMT>(out, A, pitch, W, N, x, z, mode, grid, hs); break;
        case kS2R4Y4: spmv_t<Geom2D<4, 4>, FMT>(out, A, pitch, W, N, x, z, mode, grid, hs); break;
        case kS2R4Y8: spmv_t<Geom2D<4, 8>, FMT>(out, A, pitch, W, N, x, z, mode, grid, hs); break;
        case kS2R4Y16: spmv_t<Geom2D<4, 16>, FMT>(out, A, pitch, W, N, x, z, mode, grid, hs); break;
        case kS2R8Y8: spmv_t<Geom2D<8, 8>, FMT>(out, A, pitch, W, N, x, z, mode, grid, hs); break;
        default: spmv_t<Geom2D<8, 16>, FMT>(out, A, pitch, W, N, x, z, mode, grid, hs); break;
    }
}

// SpMV-only shapes beyond the shared table: the LDS-free FP32 kernel
// (direct = 1) and the long-run FP32 kernel k_spmv_seg (direct = 2; rows per
// thread 4 G: G runs of 4 pixels).
enum { kSpD4Y4 = kNumShapes, kSpD4Y8, kSpD2Y8, kSpD2Y16, kSpSeg2, kSpSeg3, kSpSeg4, kNumSpmvShapes };
static int spmv_direct_kind(int sh) { return sh < kNumShapes ? 0 : (sh < kSpSeg2 ? 1 : 2); }
static int spmv_R(int sh) {
    if (sh < kNumShapes) return kShapeR[sh];
    if (sh >= kSpSeg2) return 4 * (sh - kSpSeg2 + 2);
    return sh <= kSpD4Y8 ? 4 : 2;
}
static int spmv_TY(int sh) {
    if (sh < kNumShapes) return kShapeTY[sh];
    if (sh >= kSpSeg2) return 4;
    return sh == kSpD4Y4 ? 4 : (sh == kSpD2Y16 ? 16 : 8);
}
static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
static bool spmv_shape_ok(int sh, int fmt, int W, long N, const float* x, size_t pitch = 0,
                          const float* out = nullptr, const float* z = nullptr) {
    if (sh < kNumShapes) return shape_ok(sh, fmt, W, N);
    if (sh >= kSpSeg2)
        return fmt == kF32 && W >= 4 && N % W == 0 && W % 4 == 0 && (double)pitch * 100.0 < 2147483648.0 &&
               aligned16(x) && aligned16(out) && aligned16(z);
    const int R = spmv_R(sh);
    return fmt == kF32 && W >= 8 && N % W == 0 && W % R == 0 &&
           (reinterpret_cast<uintptr_t>(x) & (uintptr_t)(4 * R - 1)) == 0;
}
static long spmv_tiles(int sh, int W, long N) {
    if (sh < kNumShapes) return shape_tiles(sh, W, N);
    const int TX = sh >= kSpSeg2 ? 64 * spmv_R(sh) : kBlock / spmv_TY(sh) * spmv_R(sh);
    return (long)((W + TX - 1) / TX) * ((N / W + spmv_TY(sh) - 1) / spmv_TY(sh));
}

void spmv_shape_info(int sh, int* R, int* TY, int* direct) {
    *R = spmv_R(sh);
    *TY = spmv_TY(sh);
    *direct = spmv_direct_kind(sh);
}

// y = A x (+- z) for the Levin structure with a given tile shape.
static gm_status spmv_shape_launch(float* out, const void* A, size_t pitch, int fmt, int W, int N, const float* x,
                                   const float* z, int mode, int shape, int per_cu, gm_stream_t s) {
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    if (N == 0) return GM_OK;
    if (!spmv_shape_ok(shape, fmt, W, N, x, pitch, out, z))
        return fail(GM_ERR_INVALID, "spmv: tile shape %d unusable here", shape);
    const int grid = grid_for_tiles(d, spmv_tiles(shape, W, N), per_cu);
    hipStream_t hs = (hipStream_t)s;
    const float* Af = reinterpret_cast<const float*>(A);
    if (shape == kSpSeg2) spmv_seg_t<2>(out, Af, pitch, W, N, x, z, mode, grid, hs);
    else if (shape == kSpSeg3) spmv_seg_t<3>(out, Af, pitch, W, N, x, z, mode, grid, hs);
    else if (shape == kSpSeg4) spmv_seg_t<4>(out, Af, pitch, W, N, x, z, mode, grid, hs);
    else if (shape == kSpD4Y4) spmv_direct_t<4, 4>(out, Af, pitch, W, N, x, z, mode, grid, hs);
    else if (shape == kSpD4Y8) spmv_direct_t<4, 8>(out, Af, pitch, W, N, x, z, mode, grid, hs);
    else if (shape == kSpD2Y8) spmv_direct_t<2, 8>(out, Af, pitch, W, N, x, z, mode, grid, hs);
    else if (shape == kSpD2Y16) spmv_direct_t<2, 16>(out, Af, pitch, W, N, x, z, mode, grid, hs);
    else if (fmt == kF32) spmv_shape_t<kF32>(out, A, pitch, W, N, x, z, mode, shape, grid, hs);
    else if (fmt == kF16) spmv_shape_t<kF16>(out, A, pitch, W, N, x, z, mode, shape, grid, hs);
    else spmv_shape_t<kBF16>(out, A, pitch, W, N, x, z, mode, shape, grid, hs);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

struct SpTuneKey {
    int fmt, W;
    long N;
    bool operator<(const SpTuneKey& o) const {
        return fmt != o.fmt ? fmt < o.fmt : (W != o.W ? W < o.W : N < o.N);
    }
};
static std::mutex g_sp_mu;
static std::map<SpTuneKey, std::pair<int, int>> g_sp_tune;

void spmv_tune_get(int fmt, int W, long N, int* shape, int* per_cu) {
    {
        std::lock_guard<std::mutex> g(g_sp_mu);
        auto it = g_sp_tune.find(SpTuneKey{fmt, W, N});
        if (it != g_sp_tune.end()) {
            *shape = it->second.first;
            *per_cu = it->second.second;
            return;
        }
    }
    // defaults: the shapes the tuner picks for the 3840x2160 Levin matrix (FP32:
    // the long-run form with 768-pixel tiles where W % 4 == 0; the launch falls
    // back to 256 x 4 tiles when its vectors are not 16-byte aligned)
    if (fmt == kF32 && W >= 4 && N % W == 0 && W % 4 == 0) {
        *shape = kSpSeg3;
        *per_cu = 6;
        return;
    }
    const int pref = fmt == kF32 ? kS2R4Y4 : kS2R8Y8;
    *shape = shape_ok(pref, fmt, W, N) ? pref : (fmt == kF32 ? kS1R4 : kS1R8);
    *per_cu = fmt == kF32 ? 8 : 4;
}

static thread_local std::string g_spmv_report;
const char* spmv_report() { return g_spmv_report.c_str(); }

gm_status spmv_autotune(const void* A, size_t pitch, int fmt, int W, int N, float* out, const float* x, int reps,
                        gm_stream_t s, int* best_shape, int* best_pc, double* best_us) {
    g_spmv_report.clear();
    const int pcs[] = {1, 2, 3, 4, 6, 8};
    hipEvent_t e0, e1;
    GM_HIP(hipEventCreate(&e0));
    GM_HIP(hipEventCreate(&e1));
    double best = 1e30;
    int bs = kS1R4, bp = 4;
    gm_status st = GM_OK;
    for (int sh = 0; sh < kNumSpmvShapes && st == GM_OK; ++sh) {
        if (!spmv_shape_ok(sh, fmt, W, N, x, pitch, out)) continue;
        for (int pc : pcs) {
            double tot = 0;
            for (int k = 0; k < reps + 1 && st == GM_OK; ++k) {
                GM_HIP(hipEventRecord(e0, (hipStream_t)s));
                st = spmv_shape_launch(out, A, pitch, fmt, W, N, x, nullptr, 0, sh, pc, s);
                GM_HIP(hipEventRecord(e1, (hipStream_t)s));
                GM_HIP(hipEventSynchronize(e1));
                float ms = 0;
                GM_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (k > 0) tot += ms;
            }
            const double us = tot * 1e3 / reps;
            char line[128];
            const int dk = spmv_direct_kind(sh);
            snprintf(line, sizeof(line), "spmv %d %d %d%s %d %.2f\n", sh, spmv_R(sh), spmv_TY(sh),
                     dk == 1 ? " direct" : (dk == 2 ? " runs" : ""), pc, us);
            g_spmv_report += line;
            if (us < best) {
                best = us;
                bs = sh;
                bp = pc;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != GM_OK) return st;
    {
        std::lock_guard<std::mutex> g(g_sp_mu);
        g_sp_tune[SpTuneKey{fmt, W, (long)N}] = std::make_pair(bs, bp);
    }
    *best_shape = bs;
    *best_pc = bp;
    *best_us = best;
    return GM_OK;
}

// Pin the plain product's shape for one matrix shape (tuner-table override).
gm_status spmv_tune_set(int W, int N, int R, int TY, int direct, int per_cu) {
    if (per_cu < 1 || per_cu > 16) return fail(GM_ERR_INVALID, "gm_spmv_tune_set: wg_per_cu 1..16");
    int sh = -1;
    for (int k = 0; k < kNumSpmvShapes; ++k)
        if (spmv_R(k) == R && spmv_TY(k) == TY && spmv_direct_kind(k) == direct) sh = k;
    if (sh < 0 || (sh < kNumShapes ? !shape_ok(sh, kF32, W, N)
                                   : !(W >= 8 && N % W == 0 && W % (sh >= kSpSeg2 ? 4 : R) == 0)))
        return fail(GM_ERR_INVALID, "gm_spmv_tune_set: unsupported shape");
    std::lock_guard<std::mutex> g(g_sp_mu);
    g_sp_tune[SpTuneKey{kF32, W, (long)N}] = std::make_pair(sh, per_cu);
    return GM_OK;
}

__global__ __launch_bounds__(kBlock) void k_jacobi_generic(float* __restrict__ xx, const float* __restrict__ x,
                                                           const float* __restrict__ a, size_t pitch, BandArgs ba,
                                                           int rows, const float* __restrict__ b, float omega) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < rows; i += gridDim.x * kBlock) {
        float s = b[i];
        for (int k = 0; k < ba.nb; ++k) s = fmaf(-a[(size_t)i + (size_t)k * pitch], x[(long long)i + ba.off[k]], s);
        const float xi = x[i];
        const float d = a[(size_t)i + (size_t)ba.center * pitch];
        const float v = xi + s / d;
        xx[i] = omega * v + (1.f - omega) * xi;
    }
}

bool tiled_ok(const BandPlan& p, const void* a, size_t pitch) {
    return p.levin25 && (pitch % 8) == 0 && ((reinterpret_cast<uintptr_t>(a) & 15u) == 0);
}

// y = A x (+- z) on stream s; chooses the tiled kernel for the Levin structure.
gm_status spmv_launch(float* out, const BandedMatrix* A, const BandPlan& p, const float* x, const float* z,
                      int mode, gm_stream_t s) {
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    const int N = A->rows;
    if (N == 0) return GM_OK;
    hipStream_t hs = (hipStream_t)s;
    if (tiled_ok(p, A->a, A->apitch)) {
        int shape, pc;
        spmv_tune_get(kF32, p.W, N, &shape, &pc);
        if (!spmv_shape_ok(shape, kF32, p.W, N, x, A->apitch, out, z)) {  // e.g. an unaligned x
            shape = shape_ok(kS2R4Y4, kF32, p.W, N) ? kS2R4Y4 : kS1R4;
            pc = 8;
        }
        return spmv_shape_launch(out, A->a, A->apitch, kF32, p.W, N, x, z, mode, shape, pc, s);
    } else {
        const int g = elementwise_grid(d, N, 1);
        if (mode == 0)
            hipLaunchKernelGGL(k_bm_generic<0>, dim3(g), dim3(kBlock), 0, hs, out, A->a, A->apitch, p.args, N, x, z);
        else if (mode == 1)
            hipLaunchKernelGGL(k_bm_generic<1>, dim3(g), dim3(kBlock), 0, hs, out, A->a, A->apitch, p.args, N, x, z);
        else
            hipLaunchKernelGGL(k_bm_generic<2>, dim3(g), dim3(kBlock), 0, hs, out, A->a, A->apitch, p.args, N, x, z);
    }
    GM_LAUNCH_CHECK();
    return GM_OK;
}

}  // namespace gm

using namespace gm;

extern "C" {

gm_status gm_bm_ax(float* b, const BandedMatrix* A, const float* x, gm_stream_t s) {
    BandPlan p;
    gm_status st = get_plan(A, &p);
    if (st != GM_OK) return st;
    if (A->rows > 0 && (!b || !x)) return fail(GM_ERR_INVALID, "gm_bm_ax: null vector");
    return spmv_launch(b, A, p, x, nullptr, 0, s);
}

gm_status gm_bm_axpy(float* b, const BandedMatrix* A, const float* x, const float* y, int add, gm_stream_t s) {
    BandPlan p;
    gm_status st = get_plan(A, &p);
    if (st != GM_OK) return st;
    if (A->rows > 0 && (!b || !x || !y)) return fail(GM_ERR_INVALID, "gm_bm_axpy: null vector");
    return spmv_launch(b, A, p, x, y, add ? 1 : 2, s);
}

gm_status gm_bm_jacobi_sweep(float* xx, const float* x, const BandedMatrix* A, const float* b, float omega,
                             gm_stream_t s) {
    BandPlan p;
    gm_status st = get_plan(A, &p);
    if (st != GM_OK) return st;
    if (p.args.center < 0) return fail(GM_ERR_INVALID, "matrix has no band with offset 0");
    if (A->rows > 0 && (!xx || !x || !b)) return fail(GM_ERR_INVALID, "gm_bm_jacobi_sweep: null vector");
    if (A->rows == 0) return GM_OK;
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    hipLaunchKernelGGL(k_jacobi_generic, dim3(elementwise_grid(d, A->rows, 1)), dim3(kBlock), 0, (hipStream_t)s,
                       xx, x, A->a, A->apitch, p.args, A->rows, b, omega);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

}  // extern "C"
