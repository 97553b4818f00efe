// gm_spmv.hip -- banded matrix-vector products (BandedMatrix.cu:61-207 roles).
//
// Two kernels compute y = A x (or A x +- y):
//   * k_spmv25: the Levin 5x5-band structure (offsets dy*W+dx).  A tile of
//     kBlock*R rows stages its x window (5 image-row segments of ROWS+4 floats)
//     in LDS with coalesced loads, then each thread streams the 25 diagonals
//     of its R consecutive rows with 16-byte non-temporal loads.  Persistent,
//     XCD-aware grid.  HBM-bound: N*(25*s_A + 8) algorithmic bytes.
//   * k_bm_generic: any band set (<= 64 bands), one row per thread, x read
//     directly at i + off_k (caller's zero padding, as in the reference).
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

template <int R, int FMT, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv25(float* __restrict__ out, const void* __restrict__ A, size_t pitch,
                                                   int W, int N, const float* __restrict__ x,
                                                   const float* __restrict__ z) {
    __shared__ __attribute__((aligned(16))) float lds[Tile<R>::LDS_FLOATS];
    const int ntiles = (N + Tile<R>::ROWS - 1) / Tile<R>::ROWS;
    const TileRange tr = xcd_tiles(ntiles);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int i0 = t * Tile<R>::ROWS;
        const int ib = i0 + R * threadIdx.x;
        Bands<R, FMT> av;
        if (ib < N) load_bands<R, FMT, kDiagStored>(A, pitch, ib, av);
        __syncthreads();
        stage5<R>(lds, i0, W, N, [&](int g) { return x[g]; });
        __syncthreads();
        if (ib < N) {
            float y[R];
            apply_bands<R, FMT, kDiagStored>(av, lds + R * threadIdx.x, y);
            if (MODE != 0) {
                float zz[R];
                load_rows<R>(z, ib, N, zz);
#pragma unroll
                for (int j = 0; j < R; ++j) y[j] = MODE == 1 ? y[j] + zz[j] : y[j] - zz[j];
            }
            store_rows<R>(out, ib, N, y);
        }
    }
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
        constexpr int R = 4;
        const long ntiles = (N + Tile<R>::ROWS - 1) / Tile<R>::ROWS;
        const int g = grid_for_tiles(d, ntiles, 4);
        if (mode == 0)
            hipLaunchKernelGGL((k_spmv25<R, kF32, 0>), dim3(g), dim3(kBlock), 0, hs, out, A->a, A->apitch, p.W, N, x, z);
        else if (mode == 1)
            hipLaunchKernelGGL((k_spmv25<R, kF32, 1>), dim3(g), dim3(kBlock), 0, hs, out, A->a, A->apitch, p.W, N, x, z);
        else
            hipLaunchKernelGGL((k_spmv25<R, kF32, 2>), dim3(g), dim3(kBlock), 0, hs, out, A->a, A->apitch, p.W, N, x, z);
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

// 16-bit / FP32 tiled product on a raw band array (benchmarks and the mixed
// solver's operator checks).  fmt: kF32, kF16, kBF16; R = 4 (FP32) or 8 (16-bit).
gm_status spmv25_raw(float* out, const void* A, size_t pitch, int fmt, int W, int N, const float* x, int per_cu,
                     gm_stream_t s) {
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    if (N == 0) return GM_OK;
    hipStream_t hs = (hipStream_t)s;
    if (fmt == kF32) {
        const long ntiles = (N + Tile<4>::ROWS - 1) / Tile<4>::ROWS;
        hipLaunchKernelGGL((k_spmv25<4, kF32, 0>), dim3(grid_for_tiles(d, ntiles, per_cu)), dim3(kBlock), 0, hs, out,
                           A, pitch, W, N, x, (const float*)nullptr);
    } else {
        const long ntiles = (N + Tile<8>::ROWS - 1) / Tile<8>::ROWS;
        const int g = grid_for_tiles(d, ntiles, per_cu);
        if (fmt == kF16)
            hipLaunchKernelGGL((k_spmv25<8, kF16, 0>), dim3(g), dim3(kBlock), 0, hs, out, A, pitch, W, N, x,
                               (const float*)nullptr);
        else
            hipLaunchKernelGGL((k_spmv25<8, kBF16, 0>), dim3(g), dim3(kBlock), 0, hs, out, A, pitch, W, N, x,
                               (const float*)nullptr);
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
