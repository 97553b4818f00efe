// gm_internal.h -- shared device helpers and host plumbing of libgpumatting.
//
// Layout conventions (all kernels):
//   * banded matrix: band k at a + k*apitch, row i at +i (diagonal-major)
//   * vectors: FP32, row i at v[i]; kernels that stage a tile in LDS never read
//     outside [0, N) (out-of-range neighbours are zero), generic kernels read
//     v[i + off] and rely on the caller's zero padding like the reference
//   * Levin structure: 25 bands, band k = 5*(dy+2) + (dx+2) at offset dy*W + dx
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "gpumatting.h"
#include "gm_half.h"

namespace gm {

constexpr int kBlock = 256;     // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kMaxBands = 64;
constexpr int kNumXcd = 8;      // MI355X: 8 XCDs, workgroups dispatched round-robin over them
constexpr int kMaxGrid = 4096;  // upper bound of persistent grids (partials arrays are sized by it)

// ---------------------------------------------------------------------------
// error reporting
void set_error(const char* fmt, ...);
gm_status fail(gm_status st, const char* fmt, ...);
gm_status hip_fail(hipError_t e, const char* what, const char* file, int line);

#define GM_HIP(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) return ::gm::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define GM_LAUNCH_CHECK() GM_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// band offsets handed to kernels by value
struct BandArgs {
    int nb;
    int center;          // index of the band with offset 0 (-1 if none)
    int off[kMaxBands];
};

// Host-side description of a device matrix's band structure.
struct BandPlan {
    BandArgs args;
    bool levin25 = false;  // offsets are dy*W+dx, dy,dx in [-2,2], in that order
    int W = 0;             // image width when levin25
    int maxoff = 0;        // max |offset|
};

// Looks up (or builds, with one synchronous D2H copy of the offsets) the plan
// of a device matrix.
gm_status get_plan(const BandedMatrix* A, BandPlan* out);
void register_plan(const int* dev_bands, int nbands, const int* host_bands);
void unregister_plan(const int* dev_bands);
void make_plan(const int* host_bands, int nbands, BandPlan* p);

// ---------------------------------------------------------------------------
// device info
struct DevInfo {
    int device = -1;
    int cus = 256;
};
gm_status dev_info(DevInfo* d);

// Persistent grid size for a tiled kernel: a multiple of kNumXcd, at most
// `per_cu` workgroups per CU and never more than the tile count rounded up.
int grid_for_tiles(const DevInfo& d, long ntiles, int per_cu);

// Per-(device, stream) reduction scratch: 4*kMaxGrid partial sums + counter.
struct RedScratch {
    double* partials = nullptr;
    unsigned int* counter = nullptr;
};
gm_status red_scratch(gm_stream_t s, RedScratch* out);
void release_red_scratch();
// Grid for grid-stride elementwise kernels.
int elementwise_grid(const DevInfo& d, long n, int per_thread);

// ---------------------------------------------------------------------------
// device scalar block driving the iterative solvers
// Periodic residual replacement (r := the FP64 true residual) keeps the
// search direction unless the recursion has drifted: it restarts (beta = 0)
// when ||r_true - r_rec|| > 50 % of ||r_true||, or > 10 % while the true
// residual norm has not halved since the previous replacement.  (A recursion
// run with an operator rounded to 16 bits drifts by 10-30 % per 50 iterations
// and still converges fastest with its direction kept while it makes
// progress; keeping it through poor progress has diverged.)
constexpr double kDrift2 = 0.01;
constexpr double kDriftMax2 = 0.25;
constexpr double kProgress2 = 0.25;

// Scalars::done codes
enum { kRun = 0, kConverged = 1, kBreakdown = 2, kNonFinite = 3, kMaxIt = 4 };

// Hand-off from the CG step kernel to the update kernel (see Scalars).
struct Hand {
    double rz, rz_prev, rr;
    float beta;
    int iters;
    int mode;    // kHandRun: update; kHandStop: publish this state and stop; kHandIdle: nothing to do
    int done, verify;
    int npart;   // partial sums the step kernel left (one per workgroup)
};
enum { kHandRun = 0, kHandStop = 1, kHandIdle = 2 };

// Hand-off between launches of the fused CG iteration (one slot per launch
// parity: a launch reads slot seq & 1 and writes slot (seq + 1) & 1).
struct FHand {
    double rr_prev;  // <r,r> of the residual before the one the state holds
    float beta;      // beta of the newest direction
    int iters;       // iterations completed (updates of r)
    int npart;       // workgroup partial sums left
    int xpend;       // 1: the launch deferred its update of x (x lags r by one iteration);
                     //    the next launch applies xa s p_prev first
    float xa;        // alpha of that deferred update
};

// The two CG kernels have no grid-wide reduction of their own: each
// workgroup stores its partial sums and the NEXT kernel folds them (every
// workgroup, in the same fixed order, so all agree bit for bit).  The kernel
// boundary orders the stores before the loads, so the iteration needs no
// device-scope fence or atomic (each costs an L2 write-back on gfx950, and they
// serialise per XCD).  To keep every field read-only while a kernel runs, the
// step kernel reads the Scalars and writes only `hand`; the update kernel
// reads only `hand` and writes the rest; `pend` marks update-kernel partials
// not yet folded into rz / rr.
struct Scalars {
    double rz;       // <r, z> of the current residual (z = M r)
    double rz_prev;  // previous <r, z> (kept for residual replacement)
    double rr;       // squared residual norm in the stopping norm
    double pAp;      // <p, A p>
    double bb;       // <b, b>
    double tol2;     // stop when rr <= tol2 (tol^2 * bb)
    double g0, g1;   // scratch scalars (gradient descent; g0: last true ||r||^2, g1: at the last verification)
    double rr_rep;   // true ||r||^2 at the start or the last replacement
    float alpha;     // step length
    float beta;      // direction update factor
    int done;        // 0 running, 1 converged, 2 breakdown, 3 non-finite, 4 max_iter
    int verify;      // 1: recursive residual says converged, awaiting the FP64 check
    int iters;       // iterations completed
    int maxit;
    int replacements;
    int stall;       // verifications in a row that failed to reduce the true residual 4x
    int xzero;       // k_bnorm found the initial guess all 0 (its residual is -b: no product)
    unsigned int counter;  // arrival counter of the last-block reduction
    int pend;        // 1: partial sums of <r,z>, ||r||^2 of the last update await folding
    int npart;       // number of those partial sums
    int fresh;       // batched solver: the residual was formed from scratch; the next fused launch starts from it
    Hand hand;
    FHand fh[2];
};

// ---------------------------------------------------------------------------
// device helpers

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double load_agent(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two-stage deterministic grid reduction of NQ double partial sums.  Every
// workgroup of the grid must call it exactly once.  Workgroup partials go to
// partials[q * gridDim.x + blockIdx.x]; the last workgroup to arrive sums them
// in index order (a fixed tree, independent of arrival order), calls
// fin(totals) on thread 0 and re-arms the counter.
template <int NQ, class Fin>
__device__ __forceinline__ void grid_reduce(double (&v)[NQ], double* partials, unsigned int* counter,
                                            Fin fin) {
    __shared__ double red[NQ][kWaves];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = wave_sum(v[q]);
        if (lane == 0) red[q][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) s += red[q][w];
            partials[(size_t)q * gridDim.x + blockIdx.x] = s;
        }
        __threadfence();
        unsigned int prev = atomicAdd(counter, 1u);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double tot[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = 0;
        for (int i = threadIdx.x; i < (int)gridDim.x; i += kBlock)
            s += load_agent(partials + (size_t)q * gridDim.x + i);
        s = wave_sum(s);
        if (lane == 0) red[q][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) s += red[q][w];
            tot[q] = s;
        }
        fin(tot);
        __threadfence();
        *counter = 0u;
    }
}

// Block-level sum reduction to thread 0 only (fixed order).
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kWaves; ++w) s += red[w];
    return s;
}

// Publish a hand-off state into the Scalars (the one workgroup of the update kernel that does so).
__device__ __forceinline__ void publish(Scalars* sc, const Hand& h, int done, int verify, int iters, int pend,
                                        int npart) {
    sc->rz = h.rz;
    sc->rz_prev = h.rz_prev;
    sc->rr = h.rr;
    sc->beta = h.beta;
    sc->iters = iters;
    sc->done = done;
    sc->verify = verify;
    sc->pend = pend;
    sc->npart = npart;
}

// Replace the residual by the true one (<r,z> = rz, unscaled ||r||^2 = rr;
// rsq = ||r||^2 in the space r is kept in, drift2 = ||r_true - r_rec||^2 in
// that space, periodic replacements only).  start: a fresh start (beta = 0,
// not counted).  Then the convergence / stagnation test.
__device__ __forceinline__ void replace_residual(Scalars* s, double rz, double rr, bool start, double drift2,
                                                 double rsq) {
    if (!isfinite(rz) || !isfinite(rr)) {
        s->done = kNonFinite;
        return;
    }
    if (start) {
        s->beta = 0.f;
    } else if (s->verify) {
        // the recursive residual reached the tolerance but the true one did
        // not: restart from the true residual rather than keep a direction
        // built on the drifted one
        s->beta = 0.f;
        s->replacements += 1;
        if (s->g1 > 0 && rr > 0.25 * s->g1) s->stall += 1;
        else s->stall = 0;
        s->g1 = rr;
    } else {
        const bool drifted = drift2 > kDriftMax2 * rsq || (drift2 > kDrift2 * rsq && rr > kProgress2 * s->rr_rep);
        // the <r,z> the last step used is rz while its update's partials are
        // pending (superseded by this residual), else rz_prev
        const double rzp = s->pend ? s->rz : s->rz_prev;
        s->beta = drifted ? 0.f : (rzp > 0 ? (float)(rz / rzp) : 0.f);
        s->replacements += 1;
    }
    s->rr_rep = rr;
    s->rz_prev = s->pend ? s->rz : s->rz_prev;
    s->rz = rz;
    s->rr = rr;
    s->verify = 0;
    s->pend = 0;
    if (s->tol2 >= 0 && rr <= s->tol2) s->done = kConverged;
    else if (s->iters >= s->maxit || s->stall >= 3) s->done = kMaxIt;  // max_iter or stagnation
}

// Deferred partial sums: part[q * kMaxGrid + i], i < n.  Every thread of the
// workgroup gets the totals; the summation order is fixed, so every
// workgroup of a grid computes bit-identical totals.
template <int NQ>
__device__ __forceinline__ void fold_all(const double* __restrict__ part, int n, double (&out)[NQ]) {
    __shared__ double fred[NQ][kWaves];
    __shared__ double ftot[NQ];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // every quantity's load of one index is issued before any is summed (one
    // memory latency per kBlock indices, not per quantity); each sum still runs
    // over the indices in ascending order
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) s[q] = 0;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        double v[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[q] = part[(size_t)q * kMaxGrid + i];
#pragma unroll
        for (int q = 0; q < NQ; ++q) s[q] += v[q];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double t = wave_sum(s[q]);
        if (lane == 0) fred[q][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += fred[threadIdx.x][w];
        ftot[threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = ftot[q];
}

// Store one workgroup's partial sums for the next kernel to fold.
template <int NQ>
__device__ __forceinline__ void store_partials(double (&v)[NQ], double* part) {
    __shared__ double sred[NQ][kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double s = wave_sum(v[q]);
        if (lane == 0) sred[q][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += sred[threadIdx.x][w];
        part[(size_t)threadIdx.x * kMaxGrid + blockIdx.x] = s;
    }
}

// XCD-aware tile schedule for a persistent grid (gridDim.x a multiple of
// kNumXcd).  Workgroup b runs on XCD b % 8; the tiles are cut into 8
// contiguous chunks, one per XCD, so vertically neighbouring tiles (which share
// x halo rows) are processed by the same XCD and meet in its L2.
struct TileRange {
    int begin, end, step;
};
__device__ __forceinline__ TileRange xcd_tiles(int ntiles) {
    const int G = gridDim.x, b = blockIdx.x;
    const int x = b % kNumXcd, l = b / kNumXcd, per = G / kNumXcd;
    const int chunk = (ntiles + kNumXcd - 1) / kNumXcd;
    TileRange t;
    t.begin = x * chunk + l;
    t.end = min(ntiles, (x + 1) * chunk);
    t.step = per;
    return t;
}

// Raw buffer loads with the hardware's range check: an element outside
// [0, n) -- a negative index wraps past n -- reads as 0 without touching
// memory, so window staging needs no branch per cell.  (gfx9 buffer resource:
// stride 0, num_records in bytes, dword 3 = 0x00020000.)
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, long long n_floats) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)(n_floats * 4), 0x00020000);
}
// element i (i may be negative or >= n; then 0); `skip` forces 0
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, long long i, bool skip = false) {
    const unsigned off = skip ? 0xFFFFFFFFu : (unsigned)i * 4u;
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
// elements i .. i + 3 := v (i and n multiples of 4); `skip` drops the store
// (an out-of-range buffer store writes nothing).  Kernels that store this
// way issue the same memory instructions on every path, so the compiler can
// wait for a load without waiting for the stores issued after it.
__device__ __forceinline__ void buf_st4(__amdgpu_buffer_rsrc_t r, long long i, f32x4 v, bool skip = false) {
    const unsigned off = skip ? 0x80000000u : (unsigned)i * 4u;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, (int)off, 0, 0);
}
// elements i, i + 1 := v (i and n even); `skip` drops the store
__device__ __forceinline__ void buf_st2(__amdgpu_buffer_rsrc_t r, long long i, f32x2 v, bool skip = false) {
    const unsigned off = skip ? 0x80000000u : (unsigned)i * 4u;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)off, 0, 0);
}
// elements i, i + 1 (i and n even: both in range or both 0); `skip` forces 0
__device__ __forceinline__ f32x2 buf_ld2(__amdgpu_buffer_rsrc_t r, long long i, bool skip = false) {
    const unsigned off = skip ? 0x80000000u : (unsigned)i * 4u;
    return __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
// cache policy of a buffer load that streams data read once (the NT bit)
constexpr int kNonTemporal = 2;
// elements i .. i + 3 (i and n multiples of 4, so the four are all in range or
// all 0); `skip` forces 0 with an offset far past any vector of < 2^28 floats
__device__ __forceinline__ f32x4 buf_ld4(__amdgpu_buffer_rsrc_t r, long long i, bool skip = false) {
    const unsigned off = skip ? 0x80000000u : (unsigned)i * 4u;
    return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}

typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Storage formats of matrix entries.
enum Fmt { kF32 = 0, kF16 = 1, kBF16 = 2 };

__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
    return __uint_as_float(((unsigned int)h) << 16);
}
// f32_to_f16_rne / f32_to_bf16_rne: gm_half.h

// R consecutive matrix entries of one band, kept packed in registers until use
// (16-bit formats hold two entries per dword).
template <int R, int FMT>
struct Raw;
template <int R>
struct Raw<R, kF32> {
    static constexpr int V = R >= 4 ? 4 : R;  // floats per load
    typedef float vec __attribute__((ext_vector_type(V)));
    vec d[R / V];
    __device__ __forceinline__ float get(int j) const { return d[j / V][j % V]; }
};
template <int R>
struct Raw<R, kF16> {
    unsigned int d[R / 2];
    __device__ __forceinline__ float get(int j) const {
        unsigned short h = (unsigned short)((j % 2) ? (d[j / 2] >> 16) : (d[j / 2] & 0xffffu));
        return (float)__builtin_bit_cast(_Float16, h);
    }
};
template <int R>
struct Raw<R, kBF16> {
    unsigned int d[R / 2];
    __device__ __forceinline__ float get(int j) const {
        return __uint_as_float((j % 2) ? (d[j / 2] & 0xffff0000u) : (d[j / 2] << 16));
    }
};

// Load R consecutive entries of one band starting at element i (16-byte
// aligned for FP32 / 16-bit R=8, 8-byte for 16-bit R=4) with a non-temporal
// hint: bands are streamed once per product and must not evict the vectors
// from L2 / MALL.
template <int R, int FMT>
__device__ __forceinline__ void load_band(const void* base, size_t i, Raw<R, FMT>& out) {
    if constexpr (FMT == kF32) {
        typedef typename Raw<R, kF32>::vec vec;
        constexpr int V = Raw<R, kF32>::V;
        const vec* p = reinterpret_cast<const vec*>(reinterpret_cast<const float*>(base) + i);
#pragma unroll
        for (int q = 0; q < R / V; ++q) out.d[q] = __builtin_nontemporal_load(p + q);
    } else if constexpr (R % 8 == 0) {
        const u32x4* p = reinterpret_cast<const u32x4*>(reinterpret_cast<const unsigned short*>(base) + i);
#pragma unroll
        for (int q = 0; q < R / 8; ++q) {
            u32x4 t = __builtin_nontemporal_load(p + q);
            out.d[4 * q] = t.x; out.d[4 * q + 1] = t.y; out.d[4 * q + 2] = t.z; out.d[4 * q + 3] = t.w;
        }
    } else {
        const u32x2* p = reinterpret_cast<const u32x2*>(reinterpret_cast<const unsigned short*>(base) + i);
#pragma unroll
        for (int q = 0; q < R / 4; ++q) {
            u32x2 t = __builtin_nontemporal_load(p + q);
            out.d[2 * q] = t.x; out.d[2 * q + 1] = t.y;
        }
    }
}

// Store R consecutive floats at v + i (rows >= n are skipped); vectorised when
// the whole chunk is in range and 16-byte aligned.
template <int R>
__device__ __forceinline__ void store_rows(float* v, int i, int n, const float (&y)[R]) {
    float* p = v + i;
    if (R % 4 == 0 && i + R <= n && ((reinterpret_cast<uintptr_t>(p) & 15u) == 0)) {
#pragma unroll
        for (int q = 0; q < R / 4; ++q) {
            f32x4 t = {y[4 * q], y[4 * q + 1], y[4 * q + 2], y[4 * q + 3]};
            *reinterpret_cast<f32x4*>(p + 4 * q) = t;
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j)
            if (i + j < n) p[j] = y[j];
    }
}

template <int R>
__device__ __forceinline__ void load_rows(const float* v, int i, int n, float (&y)[R]) {
    const float* p = v + i;
    if (R % 4 == 0 && i + R <= n && ((reinterpret_cast<uintptr_t>(p) & 15u) == 0)) {
#pragma unroll
        for (int q = 0; q < R / 4; ++q) {
            f32x4 t = *reinterpret_cast<const f32x4*>(p + 4 * q);
            y[4 * q] = t.x; y[4 * q + 1] = t.y; y[4 * q + 2] = t.z; y[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) y[j] = (i + j < n) ? p[j] : 0.f;
    }
}

// ---------------------------------------------------------------------------
// Levin 5x5-band tile machinery.
//
// A tile is ROWS = kBlock*R consecutive rows [i0, i0+ROWS).  Its x window is 5
// segments (dy = -2..2) of SEG = ROWS + 4 floats staged in LDS:
//   seg[s][e] = X(i0 + (s-2)*W - 2 + e),  X(g) = f(g) for 0 <= g < N, else 0.
// Thread t owns rows i0 + R*t .. + R-1 and reads seg[s][R*t .. R*t+R+3].
template <int R>
struct Tile {
    static constexpr int ROWS = kBlock * R;
    static constexpr int SEG = ROWS + 4;
    static constexpr int LDS_FLOATS = 5 * SEG;
};

template <int R, class T, class F>
__device__ __forceinline__ void stage5(T* __restrict__ lds, int i0, int W, int N, F f) {
    constexpr int SEG = Tile<R>::SEG;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        const long long gb = (long long)i0 + (long long)(s - 2) * W - 2;
        for (int e = threadIdx.x; e < SEG; e += kBlock) {
            const long long g = gb + e;
            lds[s * SEG + e] = (g >= 0 && g < N) ? f((int)g) : T(0);
        }
    }
}

// The 25 (24 with UNIT) band chunks of rows ib .. ib+R-1, loaded before the x
// window is staged so that their HBM latency overlaps the staging.
template <int R, int FMT>
struct Bands {
    Raw<R, FMT> v[25];
};

// DIAG: kDiagStored (band 12 loaded), kDiagUnit (implicit 1), kDiagCaller
// (band 12 skipped; the caller adds its own diagonal term).
enum { kDiagStored = 0, kDiagUnit = 1, kDiagCaller = 2 };

template <int R, int FMT, int DIAG>
__device__ __forceinline__ void load_bands(const void* __restrict__ A, size_t pitch, int ib, Bands<R, FMT>& b) {
#pragma unroll
    for (int k = 0; k < 25; ++k) {
        if (DIAG != kDiagStored && k == 12) continue;
        load_band<R, FMT>(A, (size_t)k * pitch + (size_t)ib, b.v[k]);
    }
}

// y[j] = sum_k A_k[i+j] * X(i+j+off_k), k = 0..24 in band order (FMA), with X
// read from the staged LDS window (band 12 per DIAG).
// STRIDE: floats between the window rows dy = -2..2 in LDS.
template <int R, int FMT, int DIAG, int STRIDE = Tile<R>::SEG>
__device__ __forceinline__ void apply_bands(const Bands<R, FMT>& av, const float* __restrict__ xs, float (&y)[R]) {
    constexpr int SEG = STRIDE;
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = 0.f;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        float xw[R + 4];
        if constexpr (R % 4 == 0) {
#pragma unroll
            for (int q = 0; q < (R + 4) / 4; ++q) {
                f32x4 t = *reinterpret_cast<const f32x4*>(xs + s * SEG + 4 * q);
                xw[4 * q] = t.x; xw[4 * q + 1] = t.y; xw[4 * q + 2] = t.z; xw[4 * q + 3] = t.w;
            }
        } else {  // R = 2: the thread's window starts 8-byte aligned
#pragma unroll
            for (int q = 0; q < (R + 4) / 2; ++q) {
                f32x2 t = *reinterpret_cast<const f32x2*>(xs + s * SEG + 2 * q);
                xw[2 * q] = t.x; xw[2 * q + 1] = t.y;
            }
        }
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
            const int k = 5 * s + dx;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                if (DIAG == kDiagUnit && k == 12)
                    y[j] += xw[j + dx];
                else if (!(DIAG == kDiagCaller && k == 12))
                    y[j] = fmaf(av.v[k].get(j), xw[j + dx], y[j]);
            }
        }
    }
}

// Loads and applies in one go (callers with nothing to overlap).
template <int R, int FMT, int DIAG>
__device__ __forceinline__ void tile_rows(const void* __restrict__ A, size_t pitch, int ib,
                                          const float* __restrict__ xs, float (&y)[R]) {
    Bands<R, FMT> av;
    load_bands<R, FMT, DIAG>(A, pitch, ib, av);
    apply_bands<R, FMT, DIAG>(av, xs, y);
}

// ---------------------------------------------------------------------------
// Tile geometries of the Levin kernels.  Both give every thread R consecutive
// rows (pixels of one image row) and an LDS window whose rows dy = -2..2 are
// STRIDE floats apart, so the same apply_bands serves both.
//
//   Geom1D<R>     : ROWS = kBlock*R consecutive matrix rows; the window is 5
//                   segments of ROWS+4 (each x element is staged by 5 tiles).
//   Geom2D<R, TY> : TX x TY pixels (TX = kBlock/TY*R); the window is the
//                   (TY+4) x (TX+4) pixel halo, ~3x fewer staged elements.
//                   Needs W % R == 0 (aligned rows); window cells are
//                   addressed by matrix-row index, so the product equals the
//                   1-D one for any band values.
template <int R_>
struct Geom1D {
    static constexpr int R = R_;
    static constexpr int ROWS = kBlock * R;
    static constexpr int STRIDE = ROWS + 4;
    static constexpr int LDS_FLOATS = 5 * STRIDE;
    static constexpr bool kIs2D = false;
    int ib;      // first row of this thread
    bool valid;  // thread has at least one row < N
    int i0;
    __device__ static int ntiles(int W, int H, int N) { return (N + ROWS - 1) / ROWS; }
    __device__ __forceinline__ void at(int t, int W, int H, int N) {
        i0 = t * ROWS;
        ib = i0 + R * (int)threadIdx.x;
        valid = ib < N;
    }
    template <class F>
    __device__ __forceinline__ void stage(float* __restrict__ lds, int W, int H, int N, F f) const {
        stage5<R>(lds, i0, W, N, f);
    }
    __device__ __forceinline__ const float* xs(const float* lds) const { return lds + R * threadIdx.x; }
    __device__ __forceinline__ float center(const float* lds, int j) const {
        return lds[2 * STRIDE + R * threadIdx.x + 2 + j];
    }
};

template <int R_, int TY_>
struct Geom2D {
    static constexpr int R = R_;
    static constexpr int TY = TY_;
    static constexpr int TPR = kBlock / TY;  // threads per tile row
    static constexpr int TX = TPR * R;
    static constexpr int STRIDE = TX + 4;
    static constexpr int LDS_FLOATS = (TY + 4) * STRIDE;
    static constexpr bool kIs2D = true;
    static_assert(kBlock % TY == 0, "tile rows must divide the workgroup");
    int ib;
    bool valid;
    int x0, y0;
    __device__ static int ntiles(int W, int H, int N) { return ((W + TX - 1) / TX) * ((H + TY - 1) / TY); }
    __device__ __forceinline__ void at(int t, int W, int H, int N) {
        const int txn = (W + TX - 1) / TX;
        x0 = (t % txn) * TX;
        y0 = (t / txn) * TY;
        const int x = x0 + R * ((int)threadIdx.x % TPR);
        const int y = y0 + (int)threadIdx.x / TPR;
        valid = x < W && y < H;
        ib = y * W + x;
    }
    template <class T, class F>
    __device__ __forceinline__ void stage(T* __restrict__ lds, int W, int H, int N, F f) const {
        constexpr int NE = (TY + 4) * STRIDE;
        for (int e = threadIdx.x; e < NE; e += kBlock) {
            const int ly = e / STRIDE, lx = e - ly * STRIDE;
            // the matrix-row index of this window cell; cells left/right of the
            // image alias the neighbouring image row exactly as the 1-D
            // product does, so both geometries compute the same A x
            const long long g = (long long)(y0 - 2 + ly) * W + (x0 - 2 + lx);
            lds[e] = (g >= 0 && g < N) ? f((int)g) : T(0);
        }
    }
    template <class T>
    __device__ __forceinline__ const T* xs(const T* lds) const {
        return lds + ((int)threadIdx.x / TPR) * STRIDE + R * ((int)threadIdx.x % TPR);
    }
    __device__ __forceinline__ float center(const float* lds, int j) const {
        return lds[((int)threadIdx.x / TPR + 2) * STRIDE + R * ((int)threadIdx.x % TPR) + 2 + j];
    }
};

// Tile shapes selectable at run time (autotuner / defaults): geometry, rows
// per thread R, tile rows TY (0 = 1-D).
enum Shape { kS1R2 = 0, kS1R4, kS1R8, kS2R4Y4, kS2R4Y8, kS2R4Y16, kS2R8Y8, kS2R8Y16, kNumShapes };
constexpr int kShapeR[kNumShapes] = {2, 4, 8, 4, 4, 4, 8, 8};
constexpr int kShapeTY[kNumShapes] = {0, 0, 0, 4, 8, 16, 8, 16};
inline bool shape_ok(int sh, int fmt, int W, long N) {
    const int R = kShapeR[sh];
    if (fmt == kF32 ? (R == 8) : (R == 2)) return false;
    if (kShapeTY[sh] == 0) return true;
    return W > 0 && N % W == 0 && W % R == 0;
}
// The 16-bit CG forms use the symmetric 12-band layout when the image width
// is a multiple of 4 (2-D tiles, 4 pixels per thread), else the full 24-band
// layout with 1-D tiles; these are the step-kernel shapes each may use.
inline bool sym16_ok(int W, long N) { return W > 0 && N % W == 0 && W % 4 == 0; }
inline bool step_shape_ok(int sh, int fmt, int W, long N) {
    if (fmt == kF32) return shape_ok(sh, fmt, W, N);
    if (sym16_ok(W, N)) return sh == kS2R4Y4 || sh == kS2R4Y8 || sh == kS2R4Y16;
    return sh == kS1R4 || sh == kS1R8;
}
inline long shape_tiles(int sh, int W, long N) {
    if (kShapeTY[sh] == 0) {
        const long rows = (long)kBlock * kShapeR[sh];
        return (N + rows - 1) / rows;
    }
    const int TX = kBlock / kShapeTY[sh] * kShapeR[sh];
    const long H = N / W;
    return (long)((W + TX - 1) / TX) * ((H + kShapeTY[sh] - 1) / kShapeTY[sh]);
}

// SpMV shape table (gm_spmv.hip) and its tuner.
void spmv_tune_get(int fmt, int W, long N, int* shape, int* per_cu);
void spmv_shape_info(int sh, int* R, int* TY, int* direct);  // rows per thread, tile rows, LDS-free form
gm_status spmv_tune_set(int W, int N, int R, int TY, int direct, int per_cu);
const char* spmv_report();  // candidates timed by the last spmv_autotune on this thread
gm_status spmv_autotune(const void* A, size_t pitch, int fmt, int W, int N, float* out, const float* x, int reps,
                        gm_stream_t s, int* best_shape, int* best_pc, double* best_us);

}  // namespace gm
