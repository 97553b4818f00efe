// gm_solve.hip -- CG, Jacobi-PCG, steepest descent and the mixed-precision
// (16-bit diagonals + FP64 residual replacement) solver on banded systems.
//
// Reference roles: cgSolve main.cu:458-526, gradSolve main.cu:406-456, the
// scalar kernels main.cu:376-404 (folded into the reductions' finalisers).
//
// One CG iteration is two kernels, all scalars stay on the device:
//   K1 k_cgstep : p_new = -z + beta*p_old is formed while the tile's x window
//                 is staged in LDS (z = r, D^-1 r, or r of the scaled system),
//                 Ap = A p_new, <p_new, Ap>   -> alpha = <r,z>/<p,Ap>
//   K2 k_cgupdate: x += alpha p, r += alpha Ap, z, <r,z>, ||r||^2
//                                            -> beta, convergence flag
// Sign convention as the reference: r = A x - b, p = -z + beta p.
// Every reduction is a deterministic two-stage sum (grid_reduce), so a solve
// is bitwise repeatable run to run.
//
// The host enqueues `check_every` iterations at a time and reads the scalar
// block of the previous batch through an event (one batch in flight), so the
// GPU never waits on the host; kernels that find the solve finished return at
// once.
#include <string.h>

#include <map>
#include <string>
#include <type_traits>
#include <utility>

#include "gm_internal.h"

namespace gm {

gm_status spmv_launch(float* out, const BandedMatrix* A, const BandPlan& p, const float* x, const float* z, int mode,
                      gm_stream_t s);
bool tiled_ok(const BandPlan& p, const void* a, size_t pitch);

enum PMode { kPlain = 0, kJacobi = 1, kScaled = 2 };


__device__ __forceinline__ bool stopped(const Scalars* sc) { return sc->done != kRun || sc->verify != 0; }

// Finaliser shared by the residual-producing kernels: record <r,z>, ||r||^2
// and test convergence.  start / !replace: beta = 0 (the next step is p = -z).
__device__ __forceinline__ void fin_residual(Scalars* sc, double rz, double rr, bool start, bool replace,
                                             double drift2 = 0.0, double rsq = 0.0) {
    replace_residual(sc, rz, rr, start || !replace, drift2, rsq);
}

// ---------------------------------------------------------------------------
// setup kernels

__global__ void k_init_scalars(Scalars* sc, int maxit) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        Scalars z = {};
        z.maxit = maxit;
        z.tol2 = -1.0;
        z.g0 = -1.0;  // no true residual measured yet
        z.counter = sc->counter;  // the reduction counter lives in the scratch block; keep it
        z.hand.mode = kHandIdle;
        *sc = z;
    }
}

// <b,b> and <x,x> partial sums per workgroup (x may be null), for k_bnorm_fin
__global__ __launch_bounds__(kBlock) void k_bnorm(const float* __restrict__ b, const float* __restrict__ x, int n,
                                                  double* __restrict__ partials) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        v[0] += (double)b[i] * (double)b[i];
        if (x) v[1] += (double)x[i] * (double)x[i];
    }
    store_partials<2>(v, partials);
}

// bb = <b,b>, tol2 = tol^2 * bb (tol <= 0: -1, never converged), xzero = 1
// when x was given and every x[i] is 0 (one workgroup)
__global__ __launch_bounds__(kBlock) void k_bnorm_fin(const double* __restrict__ partials, int npart, int has_x,
                                                      double tol, Scalars* sc) {
    double t[2];
    fold_all<2>(partials, npart, t);
    if (threadIdx.x != 0) return;
    sc->xzero = has_x && t[1] == 0.0 ? 1 : 0;
    sc->bb = t[0];
    const double bb = t[0] > 0 ? t[0] : 1.0;
    sc->tol2 = tol > 0 ? tol * tol * bb : -1.0;
}

static void launch_bnorm(const DevInfo& d, hipStream_t hs, const float* b, const float* x, int n, double tol,
                         Scalars* sc, double* partials) {
    const int g = elementwise_grid(d, n, 4);
    hipLaunchKernelGGL(k_bnorm, dim3(g), dim3(kBlock), 0, hs, b, x, n, partials);
    hipLaunchKernelGGL(k_bnorm_fin, dim3(1), dim3(kBlock), 0, hs, partials, g, x ? 1 : 0, tol, sc);
}

// ---------------------------------------------------------------------------
// tiled kernels (Levin 5x5 structure)

template <int PM>
__device__ __forceinline__ float zval(const float* __restrict__ r, const float* __restrict__ dinv, int g) {
    if constexpr (PM == kJacobi) return dinv[g] * r[g];
    else return r[g];
}

// r = A x - b (FP32), z, <r,z>, ||r||^2.  start: also the first step (beta=0).
template <int R, int PM>
__global__ __launch_bounds__(kBlock) void k_resid25(const float* __restrict__ A, size_t pitch, int W, int N,
                                                    const float* __restrict__ x, const float* __restrict__ b,
                                                    const float* __restrict__ dinv, float* __restrict__ r,
                                                    Scalars* sc, double* partials, unsigned int* counter, int start) {
    if (!start && sc->done != kRun) return;
    __shared__ __attribute__((aligned(16))) float lds[Tile<R>::LDS_FLOATS];
    const int ntiles = (N + Tile<R>::ROWS - 1) / Tile<R>::ROWS;
    const TileRange tr = xcd_tiles(ntiles);
    double v[2] = {0.0, 0.0};
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int i0 = t * Tile<R>::ROWS;
        __syncthreads();
        int nz = 0;
        stage5<R>(lds, i0, W, N, [&](int g) {
            const float v = x[g];
            nz |= v != 0.f;
            return v;
        });
        // a window of zeros (a zero initial guess) needs no band loads: A x = 0
        nz = __syncthreads_or(nz);
        const int ib = i0 + R * threadIdx.x;
        if (ib < N) {
            float y[R], bb[R], dd[R];
            if (nz) {
                tile_rows<R, kF32, kDiagStored>(A, pitch, ib, lds + R * threadIdx.x, y);
            } else {
#pragma unroll
                for (int j = 0; j < R; ++j) y[j] = 0.f;
            }
            load_rows<R>(b, ib, N, bb);
            if (PM == kJacobi) load_rows<R>(dinv, ib, N, dd);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                y[j] -= bb[j];
                if (ib + j < N) {
                    const float z = PM == kJacobi ? dd[j] * y[j] : y[j];
                    v[0] += (double)(y[j] * z);
                    v[1] += (double)(y[j] * y[j]);
                }
            }
            store_rows<R>(r, ib, N, y);
        }
    }
    grid_reduce<2>(v, partials, counter, [&](double* tt) {
        sc->g0 = -1.0;  // an FP32 residual: no FP64 true residual of this x yet
        fin_residual(sc, tt[0], tt[1], true, false);
    });
}

// Start of the CG step kernel: fold the update kernel's pending partial sums
// (<r,z>, ||r||^2 of the new residual) and run the convergence test the
// update would have run.  Returns false when the iteration stops here (block 0
// has recorded why in sc->hand); fills rz, rz_prev, rr, beta, iters otherwise.
struct StepIn {
    double rz, rz_prev, rr;
    float beta;
    int iters;
};
__device__ __forceinline__ bool step_begin(Scalars* sc, const double* __restrict__ dpb, int verify, StepIn& in) {
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (sc->done != kRun || sc->verify) {
        if (lead) sc->hand.mode = kHandIdle;
        return false;
    }
    in.rz = sc->rz;
    in.rz_prev = sc->rz_prev;
    in.rr = sc->rr;
    in.beta = sc->beta;
    in.iters = sc->iters;
    if (!sc->pend) return true;
    double t[2];
    fold_all<2>(dpb, sc->npart, t);
    int stop = kRun, ver = 0;
    if (!isfinite(t[0]) || !isfinite(t[1])) {
        stop = kNonFinite;
    } else {
        in.rz_prev = in.rz;
        in.beta = (float)(t[0] / in.rz);
        in.rz = t[0];
        in.rr = t[1];
        if (sc->tol2 >= 0 && in.rr <= sc->tol2) {
            if (verify) ver = 1;
            else stop = kConverged;
        } else if (in.iters >= sc->maxit) {
            stop = kMaxIt;
        }
    }
    if (stop == kRun && !ver) return true;
    if (lead) sc->hand = Hand{in.rz, in.rz_prev, in.rr, in.beta, in.iters, kHandStop, stop, ver, 0};
    return false;
}

// End of the CG step kernel: this workgroup's <p, Ap> partial; block 0 hands
// the iteration state to the update kernel.
__device__ __forceinline__ void step_end(Scalars* sc, double v, double* __restrict__ dpa, const StepIn& in) {
    double vv[1] = {v};
    store_partials<1>(vv, dpa);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        sc->hand = Hand{in.rz, in.rz_prev, in.rr, in.beta, in.iters, kHandRun, kRun, 0, (int)gridDim.x};
}

// K1: p_new = -z + beta p_old (staged), Ap = A p_new, <p_new, Ap>.
// G: tile geometry (Geom1D / Geom2D).
template <class G, int FMT, int PM>
__global__ __launch_bounds__(kBlock) void k_cgstep(const void* __restrict__ A, size_t pitch, int W, int H, int N,
                                                   const float* __restrict__ r, const float* __restrict__ dinv,
                                                   float* __restrict__ p0, float* __restrict__ p1,
                                                   float* __restrict__ Ap, Scalars* sc, double* __restrict__ dpa,
                                                   const double* __restrict__ dpb, int verify) {
    StepIn in;
    if (!step_begin(sc, dpb, verify, in)) return;
    constexpr int R = G::R;
    constexpr int DIAG = PM == kScaled ? kDiagUnit : kDiagStored;
    const float beta = in.beta;
    const int par = in.iters & 1;
    const float* __restrict__ pold = par ? p1 : p0;
    float* __restrict__ pnew = par ? p0 : p1;
    __shared__ __attribute__((aligned(16))) float lds[G::LDS_FLOATS];
    const TileRange tr = xcd_tiles(G::ntiles(W, H, N));
    double v[1] = {0.0};
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        G g;
        g.at(t, W, H, N);
        Bands<R, FMT> av;
        if (g.valid) load_bands<R, FMT, DIAG>(A, pitch, g.ib, av);
        __syncthreads();
        if (beta != 0.f)
            g.stage(lds, W, H, N, [&](int i) { return fmaf(beta, pold[i], -zval<PM>(r, dinv, i)); });
        else
            g.stage(lds, W, H, N, [&](int i) { return -zval<PM>(r, dinv, i); });
        __syncthreads();
        if (g.valid) {
            float y[R], pc[R];
            apply_bands<R, FMT, DIAG, G::STRIDE>(av, g.xs(lds), y);
#pragma unroll
            for (int j = 0; j < R; ++j) pc[j] = g.center(lds, j);
            store_rows<R>(pnew, g.ib, N, pc);
            store_rows<R>(Ap, g.ib, N, y);
#pragma unroll
            for (int j = 0; j < R; ++j)
                if (g.ib + j < N) v[0] += (double)(pc[j] * y[j]);
        }
    }
    step_end(sc, v[0], dpa, in);
}

// One 16-bit band word (low half of h) as float.
template <int FMT>
__device__ __forceinline__ float h2f(unsigned int h) {
    if constexpr (FMT == kF16) return (float)__builtin_bit_cast(_Float16, (unsigned short)(h & 0xffffu));
    else return __uint_as_float(h << 16);
}

// K1 on the symmetric scaled operator.  Only the 12 upper bands k = 13..24
// (offset o_k > 0) of As = S A S are stored, the unit diagonal is implicit,
// and row i's entry in the lower band 24 - k is the upper entry of row
// i - o_k, U_k[i - o_k].  Each thread loads its own rows' upper entries into
// registers and also into an LDS window of U covering the tile plus 2 image
// rows above and 2 columns on each side (window cells are addressed by
// matrix-row index, like the x window); the lower entries are read from it.
// U holds 0 for pairs outside the image or wrapping a row, so the product
// equals the full 24-band one wherever As is symmetric.  16-bit bands move
// 24 B per row instead of 48.
// FP32 upper bands (the fused FP32 iteration) take 2 cells per thread, so that
// a thread's band words, and a staged U row's 32-bit words, are as many as
// for 16-bit bands with 4 cells per thread.
template <int TY, int FMT = kF16>
struct SymGeom {
    static constexpr int R = FMT == kF32 ? 2 : 4;    // cells per thread
    static constexpr int EPW = FMT == kF32 ? 1 : 2;  // band entries per 32-bit word
    using G = Geom2D<R, TY>;
    static constexpr int TPR = G::TPR, TX = G::TX, ST = G::STRIDE;
    static constexpr int UR = TY + 2;          // staged U rows: 2 above the tile, then the tile's
    static constexpr int SW = (TX + 4) / EPW;  // 32-bit words per staged U row: TX + 4 entries
    static constexpr int XR = TY + 4;          // staged x rows
};

// Entry q of a run of band words (16-bit: two per word, low half first).
template <int FMT>
__device__ __forceinline__ float band_entry(const unsigned int* w, int q) {
    if constexpr (FMT == kF32) return __uint_as_float(w[q]);
    else return h2f<FMT>(w[q >> 1] >> (16 * (q & 1)));
}

// One tile's product from the staged windows.  xrow[q]: LDS offset of the
// thread's R + 4 x cells (columns R tx .. R tx + R + 3) in window row ty + q;
// urow[d]: LDS row of U window row ty + 2 - d.  Sums in band order, so every
// staging scheme gives bit-identical results.
template <int TY, int FMT, int URR = SymGeom<TY, FMT>::UR>
__device__ __forceinline__ void sym_tile(const float* __restrict__ lds, const int (&xrow)[5],
                                         const unsigned int (*ub)[URR][SymGeom<TY, FMT>::SW],
                                         const int (&urow)[3], int tx, float (&y)[SymGeom<TY, FMT>::R],
                                         float (&pc)[SymGeom<TY, FMT>::R]) {
    using SG = SymGeom<TY, FMT>;
    constexpr int R = SG::R, NC = R + 4, NW = NC / SG::EPW;  // window cells, and their band words
    float xw[5][NC];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        if constexpr (R == 4) {
            const f32x4 a0 = *reinterpret_cast<const f32x4*>(lds + xrow[q]);
            const f32x4 a1 = *reinterpret_cast<const f32x4*>(lds + xrow[q] + 4);
            xw[q][0] = a0.x; xw[q][1] = a0.y; xw[q][2] = a0.z; xw[q][3] = a0.w;
            xw[q][4] = a1.x; xw[q][5] = a1.y; xw[q][6] = a1.z; xw[q][7] = a1.w;
        } else {
#pragma unroll
            for (int h = 0; h < NC / 2; ++h) {
                const f32x2 a = *reinterpret_cast<const f32x2*>(lds + xrow[q] + 2 * h);
                xw[q][2 * h] = a.x;
                xw[q][2 * h + 1] = a.y;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        pc[j] = xw[2][2 + j];
        y[j] = pc[j];  // unit diagonal
    }
#pragma unroll
    for (int m = 0; m < 12; ++m) {
        const int k = 13 + m, dy = k / 5 - 2, dx = k % 5 - 2;
        // words of cells R tx .. R tx + R + 3 of window rows ty + 2 (own) and ty + 2 - dy
        unsigned int wn[NW], wl[NW];
        const unsigned int* wo = &ub[m][urow[0]][R * tx / SG::EPW];
#pragma unroll
        for (int h = 0; h < NW / 2; ++h) {
            const uint2 o = *reinterpret_cast<const uint2*>(wo + 2 * h);
            wn[2 * h] = wl[2 * h] = o.x;
            wn[2 * h + 1] = wl[2 * h + 1] = o.y;
        }
        if (dy != 0) {
            const unsigned int* wr = &ub[m][urow[dy]][R * tx / SG::EPW];
#pragma unroll
            for (int h = 0; h < NW / 2; ++h) {
                const uint2 o = *reinterpret_cast<const uint2*>(wr + 2 * h);
                wl[2 * h] = o.x;
                wl[2 * h + 1] = o.y;
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            // upper entry of this row (cell 2 + j): x at (+dy, +dx)
            y[j] = fmaf(band_entry<FMT>(wn, 2 + j), xw[2 + dy][2 + dx + j], y[j]);
            // lower entry: U_k of the row at (-dy, -dx) (cell 2 + j - dx): x there
            y[j] = fmaf(band_entry<FMT>(wl, 2 + j - dx), xw[2 - dy][2 - dx + j], y[j]);
        }
    }
}

// Tile-parallel form: tiles in row-major order, each workgroup takes every
// step-th tile of its XCD's chunk and stages the whole window.
template <int TY, int FMT>
__global__ __launch_bounds__(kBlock) void k_cgstep_sym(const unsigned short* __restrict__ U, size_t pitch, int W,
                                                       int H, int N, const float* __restrict__ r,
                                                       float* __restrict__ p0, float* __restrict__ p1,
                                                       float* __restrict__ Ap, Scalars* sc, double* __restrict__ dpa,
                                                       const double* __restrict__ dpb, int verify) {
    using SG = SymGeom<TY>;
    using G = typename SG::G;
    constexpr int R = 4, TPR = SG::TPR, ST = SG::ST, UR = SG::UR, SW = SG::SW;
    StepIn in;
    if (!step_begin(sc, dpb, verify, in)) return;
    const float beta = in.beta;
    const int par = in.iters & 1;
    const float* __restrict__ pold = par ? p1 : p0;
    float* __restrict__ pnew = par ? p0 : p1;
    __shared__ __attribute__((aligned(16))) float lds[G::LDS_FLOATS];
    __shared__ __attribute__((aligned(16))) unsigned int ub[12][UR][SW];
    const TileRange tr = xcd_tiles(G::ntiles(W, H, N));
    const int tx = (int)threadIdx.x % TPR, ty = (int)threadIdx.x / TPR;
    constexpr int NE = (TY + 4) * ST;  // x window cells
    constexpr int XE = (NE + kBlock - 1) / kBlock;
    constexpr int NH = 2 * SW + 2 * TY;  // U halo words: all of window rows 0, 1; words 0 and SW - 1 below
    constexpr int HE = (NH + kBlock - 1) / kBlock;
    // Software pipeline: the global loads of tile t + step are issued right
    // after tile t's are moved to LDS, so they are in flight while tile t is
    // computed from LDS.  One register set holds a tile's loads.
    struct Loads {
        Raw<R, FMT> u[12];  // own rows' upper bands (0 outside the image)
        unsigned int hv[HE][12];
        float xp[XE], xr[XE];  // p_old and r of the x window cells (0 outside)
    };
    auto issue = [&](int t, Loads& L) {
        G g;
        g.at(t, W, H, N);
#pragma unroll
        for (int m = 0; m < 12; ++m) {
            if (g.valid) load_band<R, FMT>(U, (size_t)(13 + m) * pitch + (size_t)g.ib, L.u[m]);
            else L.u[m].d[0] = L.u[m].d[1] = 0u;
        }
#pragma unroll
        for (int q = 0; q < HE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            int wy, ww;
            if (e < 2 * SW) {
                wy = e / SW;
                ww = e - wy * SW;
            } else {
                const int f = e - 2 * SW;
                wy = 2 + (f >> 1);
                ww = (f & 1) ? SW - 1 : 0;
            }
            const long long gi = (long long)(g.y0 - 2 + wy) * W + (g.x0 - 2 + 2 * ww);  // even: W % 4 == 0
            const bool in = e < NH && gi >= 0 && gi < N;
#pragma unroll
            for (int m = 0; m < 12; ++m)
                L.hv[q][m] = in ? *reinterpret_cast<const unsigned int*>(U + (size_t)(13 + m) * pitch + (size_t)gi) : 0u;
        }
#pragma unroll
        for (int q = 0; q < XE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            const int ly = e / ST, lx = e - ly * ST;
            // matrix-row index of the window cell (as Geom2D::stage)
            const long long gi = (long long)(g.y0 - 2 + ly) * W + (g.x0 - 2 + lx);
            const bool in = e < NE && gi >= 0 && gi < N;
            L.xp[q] = in ? pold[gi] : 0.f;
            L.xr[q] = in ? r[gi] : 0.f;
        }
    };
    int xrow[5], urow[3];
#pragma unroll
    for (int q = 0; q < 5; ++q) xrow[q] = (ty + q) * ST + R * tx;
#pragma unroll
    for (int d = 0; d < 3; ++d) urow[d] = ty + 2 - d;
    Loads L;
    if (tr.begin < tr.end) issue(tr.begin, L);
    double v[1] = {0.0};
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        G g;
        g.at(t, W, H, N);
        __syncthreads();  // the previous tile is done with lds and ub
        // own cells: window row ty + 2, halves 2 + 4 tx .. 5 + 4 tx (words 1 + 2 tx, 2 + 2 tx)
#pragma unroll
        for (int m = 0; m < 12; ++m) {
            ub[m][ty + 2][1 + 2 * tx] = L.u[m].d[0];
            ub[m][ty + 2][2 + 2 * tx] = L.u[m].d[1];
        }
#pragma unroll
        for (int q = 0; q < HE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            if (e < NH) {
                int wy, ww;
                if (e < 2 * SW) {
                    wy = e / SW;
                    ww = e - wy * SW;
                } else {
                    const int f = e - 2 * SW;
                    wy = 2 + (f >> 1);
                    ww = (f & 1) ? SW - 1 : 0;
                }
#pragma unroll
                for (int m = 0; m < 12; ++m) ub[m][wy][ww] = L.hv[q][m];
            }
        }
#pragma unroll
        for (int q = 0; q < XE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            if (e < NE) lds[e] = beta != 0.f ? fmaf(beta, L.xp[q], -L.xr[q]) : -L.xr[q];
        }
        __syncthreads();
        if (t + tr.step < tr.end) issue(t + tr.step, L);
        if (g.valid) {
            float y[R], pc[R];
            sym_tile<TY, FMT>(lds, xrow, ub, urow, tx, y, pc);
            store_rows<R>(pnew, g.ib, N, pc);
            store_rows<R>(Ap, g.ib, N, y);
#pragma unroll
            for (int j = 0; j < R; ++j)
                if (g.ib + j < N) v[0] += (double)(pc[j] * y[j]);
        }
    }
    step_end(sc, v[0], dpa, in);
}

// Column-walk form: the tiles are numbered column-major (down each TX-wide
// column of tiles) and every workgroup takes one contiguous run of them.  Going
// down a column, the window rows a tile shares with the tile above (4 x rows,
// 2 U rows) are already in LDS: the x and U windows are row rings, and each
// tile loads only its own U rows (plus the two side words of each) and the TY x
// rows below the shared ones.  The first tile of a run, and a tile starting a
// new column, also load the shared rows.  Every x and U element is then read
// about once per launch.
// Held to 168 VGPRs: 3 workgroups per CU (LDS allows 3 at 4-row tiles)
template <int TY, int FMT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_cgstep_symw(const unsigned short* __restrict__ U, size_t pitch, int W,
                                                        int H, int N, const float* __restrict__ r,
                                                        float* __restrict__ p0, float* __restrict__ p1,
                                                        float* __restrict__ Ap, Scalars* sc,
                                                        double* __restrict__ dpa, const double* __restrict__ dpb,
                                                        int verify) {
    using SG = SymGeom<TY>;
    constexpr int R = 4, TPR = SG::TPR, TX = SG::TX, ST = SG::ST, UR = SG::UR, SW = SG::SW, XR = SG::XR;
    StepIn in;
    if (!step_begin(sc, dpb, verify, in)) return;
    const float beta = in.beta;
    const int par = in.iters & 1;
    const float* __restrict__ pold = par ? p1 : p0;
    float* __restrict__ pnew = par ? p0 : p1;
    __shared__ __attribute__((aligned(16))) float lds[XR * ST];
    __shared__ __attribute__((aligned(16))) unsigned int ub[12][UR][SW];
    const int ncol = (W + TX - 1) / TX, nrow = (H + TY - 1) / TY;
    // Work: one run [ubeg, uend) of the column-major tile order per workgroup.
    // With at least one workgroup per column on each XCD, the runs form a
    // grid: each XCD takes a band of tile rows, split into s segments, and its
    // workgroup l walks column l % ncol of segment l / ncol.  Workgroups of one
    // segment then walk the same rows side by side on the same XCD, so the side
    // columns a tile reads from its neighbours' tiles come from that XCD's L2.
    // Otherwise the column-major order is cut into equal runs.
    int ubeg, uend;
    {
        const int g = gridDim.x / kNumXcd, xcd = blockIdx.x % kNumXcd, l = blockIdx.x / kNumXcd;
        if (gridDim.x % kNumXcd == 0 && g >= ncol) {
            const int sg = g / ncol, nseg = kNumXcd * sg;
            if (l < ncol * sg) {
                const int col = l % ncol, seg = xcd * sg + l / ncol;
                ubeg = col * nrow + (int)((long long)seg * nrow / nseg);
                uend = col * nrow + (int)((long long)(seg + 1) * nrow / nseg);
            } else {
                ubeg = uend = 0;
            }
        } else {
            const long long T = (long long)ncol * nrow;
            ubeg = (int)(T * blockIdx.x / gridDim.x);
            uend = (int)(T * (blockIdx.x + 1) / gridDim.x);
        }
    }
    const int tx = (int)threadIdx.x % TPR, ty = (int)threadIdx.x / TPR;
    constexpr int NX = TY * ST;  // new x cells per tile: window rows 4 .. TY + 3
    constexpr int XE = (NX + kBlock - 1) / kBlock;
    constexpr int NS = 2 * TY;  // side words of the own U rows: words 0 and SW - 1 of window rows 2 .. TY + 1
    constexpr int SE = (12 * NS + kBlock - 1) / kBlock;  // side words per thread (band e / NS, word e % NS)
    struct Loads {
        Raw<R, FMT> u[12];
        unsigned int sv[SE];
        float xp[XE], xr[XE];
    };
    const __amdgpu_buffer_rsrc_t bp = buf_rsrc(pold, N), br = buf_rsrc(r, N);
    // tile u: column u / nrow, tile row u % nrow
    auto origin = [&](int u, int& x0, int& y0) {
        x0 = (u / nrow) * TX;
        y0 = (u % nrow) * TY;
    };
    auto issue = [&](int u, Loads& L) {
        int x0, y0;
        origin(u, x0, y0);
        const int x = x0 + R * tx, y = y0 + ty;
        const bool valid = x < W && y < H;
        const size_t ib = (size_t)y * W + x;
#pragma unroll
        for (int m = 0; m < 12; ++m) {
            if (valid) load_band<R, FMT>(U, (size_t)(13 + m) * pitch + ib, L.u[m]);
            else L.u[m].d[0] = L.u[m].d[1] = 0u;
        }
#pragma unroll
        for (int q = 0; q < SE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock, m = e / NS, f = e - m * NS;
            const int wy = 2 + (f >> 1), ww = (f & 1) ? SW - 1 : 0;
            const long long gi = (long long)(y0 - 2 + wy) * W + (x0 - 2 + 2 * ww);  // even: W % 4 == 0
            const bool in = e < 12 * NS && gi >= 0 && gi < N;
            L.sv[q] = in ? *reinterpret_cast<const unsigned int*>(U + (size_t)(13 + m) * pitch + (size_t)gi) : 0u;
        }
#pragma unroll
        for (int q = 0; q < XE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            const int ly = 4 + e / ST, lx = e % ST;
            const long long gi = (long long)(y0 - 2 + ly) * W + (x0 - 2 + lx);
            L.xp[q] = buf_ld(bp, gi, e >= NX);
            L.xr[q] = buf_ld(br, gi, e >= NX);
        }
    };
    Loads L;
    if (ubeg < uend) issue(ubeg, L);
    int xtop = 0, utop = 0;  // LDS rows of window row 0 (x ring of XR rows, U ring of UR rows)
    double v[1] = {0.0};
    for (int u = ubeg; u < uend; ++u) {
        int x0, y0;
        origin(u, x0, y0);
        const bool first = u == ubeg || y0 == 0;
        __syncthreads();  // the previous tile is done with lds and ub
        if (first) {
            // the rows shared with the tile above: U window rows 0, 1 and x
            // window rows 0 .. 3, read straight into LDS
            xtop = 0;
            utop = 0;
            for (int e = (int)threadIdx.x; e < 2 * SW; e += kBlock) {
                const int wy = e / SW, ww = e - wy * SW;
                const long long gi = (long long)(y0 - 2 + wy) * W + (x0 - 2 + 2 * ww);
                const bool in = gi >= 0 && gi < N;
#pragma unroll
                for (int m = 0; m < 12; ++m)
                    ub[m][wy][ww] =
                        in ? *reinterpret_cast<const unsigned int*>(U + (size_t)(13 + m) * pitch + (size_t)gi) : 0u;
            }
            for (int e = (int)threadIdx.x; e < 4 * ST; e += kBlock) {
                const int ly = e / ST, lx = e - ly * ST;
                const long long gi = (long long)(y0 - 2 + ly) * W + (x0 - 2 + lx);
                const float xp = buf_ld(bp, gi), xr = buf_ld(br, gi);
                lds[e] = beta != 0.f ? fmaf(beta, xp, -xr) : -xr;
            }
        } else {
            // window rows TY.. of the tile above are rows 0.. of this one
            xtop = xtop + TY < XR ? xtop + TY : xtop + TY - XR;
            utop = utop + TY < UR ? utop + TY : utop + TY - UR;
        }
        // own U rows (window rows 2 .. TY + 1) and their side words
        {
            const int wr = utop + 2 + ty < UR ? utop + 2 + ty : utop + 2 + ty - UR;
#pragma unroll
            for (int m = 0; m < 12; ++m) {
                ub[m][wr][1 + 2 * tx] = L.u[m].d[0];
                ub[m][wr][2 + 2 * tx] = L.u[m].d[1];
            }
#pragma unroll
            for (int q = 0; q < SE; ++q) {
                const int e = (int)threadIdx.x + q * kBlock;
                if (e < 12 * NS) {
                    const int m = e / NS, f = e - m * NS;
                    const int wy = utop + 2 + (f >> 1), ws = wy < UR ? wy : wy - UR;
                    const int ww = (f & 1) ? SW - 1 : 0;
                    ub[m][ws][ww] = L.sv[q];
                }
            }
        }
        // x window rows 4 .. TY + 3
#pragma unroll
        for (int q = 0; q < XE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            if (e < NX) {
                const int ly = xtop + 4 + e / ST, lx = e % ST;
                const int ls = ly < XR ? ly : ly - XR;
                lds[ls * ST + lx] = beta != 0.f ? fmaf(beta, L.xp[q], -L.xr[q]) : -L.xr[q];
            }
        }
        __syncthreads();
        if (u + 1 < uend) issue(u + 1, L);
        const int x = x0 + R * tx, y = y0 + ty;
        if (x < W && y < H) {
            int xrow[5], urow[3];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const int lr = xtop + ty + q;
                xrow[q] = (lr < XR ? lr : lr - XR) * ST + R * tx;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int lr = utop + ty + 2 - d;
                urow[d] = lr < UR ? lr : lr - UR;
            }
            float yv[R], pc[R];
            sym_tile<TY, FMT>(lds, xrow, ub, urow, tx, yv, pc);
            const int ib = y * W + x;
            store_rows<R>(pnew, ib, N, pc);
            store_rows<R>(Ap, ib, N, yv);
#pragma unroll
            for (int j = 0; j < R; ++j)
                if (ib + j < N) v[0] += (double)(pc[j] * yv[j]);
        }
    }
    step_end(sc, v[0], dpa, in);
}

// Fused CG iteration on the symmetric 16-bit operator (column walk, 4-row
// tiles).  One launch finishes iteration j and makes the step of j + 1:
//   update j   x += alpha_j s p_j,  r_{j+1} = r_j + alpha_j Ap_j
//   step j+1   p_{j+1} = -r_{j+1} + beta_{j+1} p_j,  Ap_{j+1} = A p_{j+1}
// alpha_j = <r_j,r_j> / <p_j,Ap_j> uses the previous launch's sums, and so
// does beta_{j+1} = <r_{j+1},r_{j+1}> / <r_j,r_j> through
//   <r_{j+1},r_{j+1}> = <r_j,r_j> + 2 alpha_j <r_j,Ap_j> + alpha_j^2 <Ap_j,Ap_j>
// (the product expanded).  Every launch also sums the exact <r,r> of the
// residual it stores, which the next expansion starts from, so rounding does
// not accumulate across iterations.  Sums per workgroup: <p,Ap>, <r,r>,
// <r,Ap>, <Ap,Ap> in the scaled space and the unscaled ||r/s||^2 (stopping
// test), all in FP64.
// x and r follow k_cgupdate's rounding exactly.  r and Ap are read at window
// cells of neighbouring tiles while their owners write the new ones, so both
// alternate between two buffers by iteration parity, as p does.
// Window loads: each thread loads the 4 interior cells of one new window row,
// the row it outputs now (threads of tile rows 2, 3) or in the next tile
// (rows 0, 1, held in registers until then), so r_{j+1} and p_j of its own
// cells never have to be read twice.
// mode 0: continue from the previous launch (slot seq & 1); 1: first step
// after a residual from scratch (state from the Scalars, no update: the step
// j of the current residual); 2: timing (as 0 with alpha = beta = 0.5 and an
// update of x deferred by the launch before, no sums folded, no stop test);
// 3: the update of iteration j only, when the residual is replaced next (its
// step would be redone from the new residual): x and r_{j+1} of every cell,
// <r,r> and ||r/s||^2 of r_{j+1}; 4: as 0, but the update of x is deferred to
// the next launch; 5: timing of mode 4.
// Deferring x: a launch in mode 4 neither reads nor writes x (52 instead of
// 60 B per row with 16-bit bands; s is still read for ||r/s||^2, so every
// launch runs the convergence test).  The launch after it finds alpha_{j-1}
// in its hand-off slot and p_{j-1} in the buffer it writes p_{j+1} to (each
// cell is read by the thread that overwrites it, before it does), and makes
// both updates, x + alpha_{j-1} s p_{j-1} + alpha_j s p_j, with the roundings
// of two launches (64 B per row): x is bit-identical to the undeferred
// sequence.  A stop that launch decides first brings x up to date.  The host
// never defers the last launch before a fold, a replacement or the iteration
// cap.
// The loads of tile u + 1 are in flight while
// tile u is computed; two workgroups per CU (the tuner's choice: a third, or a
// second tile of loads in flight, measured slower at 4K).
// The side words of the symmetric 16-bit operator's column walk, gathered
// once per operator: for every column of TX-wide tiles, image row and upper
// band, the two 32-bit words (half pairs) just left (columns x0 - 2, x0 - 1)
// and right (x0 + TX, x0 + TX + 1) of the column, at
// E[((column * H + row) * 12 + band) * 2 + side], 0 outside the vector (the
// indexing of the window cells: rows wrap at the image edge).  A tile's 96
// side words are then 384 contiguous bytes instead of 96 cache lines that the
// neighbouring columns' tiles also read.
__global__ __launch_bounds__(kBlock) void k_uedge(const unsigned short* __restrict__ U, size_t pitch, int W, int H,
                                                  int N, int TX, unsigned int* __restrict__ E) {
    const int ncol = (W + TX - 1) / TX;
    const long long total = (long long)ncol * H * 24;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total; i += (long long)gridDim.x * kBlock) {
        const int rem = (int)(i % 24), m = rem >> 1, side = rem & 1;
        const long long cy = i / 24;
        const int y = (int)(cy % H), c = (int)(cy / H);
        const long long gi = (long long)y * W + (side ? c * TX + TX : c * TX - 2);  // even: W % 4 == 0
        E[i] = (gi >= 0 && gi < N) ? *reinterpret_cast<const unsigned int*>(U + (size_t)(13 + m) * pitch + gi) : 0u;
    }
}

// The same side words of the FP32 upper bands (128-wide columns, 2 cells per
// side): E[(((column * H + row) * 12 + band) * 2 + side) * 2 + w] =
// U_band[row * W + column * TX - 2 + w] (left) or U_band[... + TX + w]
// (right), 0 outside the vector.  U holds band 13 + m at U + m * pitch.
__global__ __launch_bounds__(kBlock) void k_uedge32(const float* __restrict__ U, size_t pitch, int W, int H, int N,
                                                    int TX, unsigned int* __restrict__ E) {
    const int ncol = (W + TX - 1) / TX;
    const long long total = (long long)ncol * H * 48;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total; i += (long long)gridDim.x * kBlock) {
        const int rem = (int)(i % 48), m = rem >> 2, side = (rem >> 1) & 1, wd = rem & 1;
        const long long cy = i / 48;
        const int y = (int)(cy % H), c = (int)(cy / H);
        const long long gi = (long long)y * W + (side ? c * TX + TX : c * TX - 2) + wd;
        E[i] = (gi >= 0 && gi < N) ? __float_as_uint(U[(size_t)m * pitch + gi]) : 0u;
    }
}

// Vectors of fewer than fused_max_n(FMT) elements: the descriptor over the 12
// upper bands spans 12 x pitch entries (24 or 48 x pitch bytes), which must
// stay below 2^31.
constexpr int kFusedMaxN = 1 << 26;
constexpr int kFused32MaxN = 1 << 25;
inline int fused_max_n(int fmt) { return fmt == kF32 ? kFused32MaxN : kFusedMaxN; }
// Words of the fused iteration's side-word array of a W x H operator.
inline long fused_edge_words(int fmt, int W, long N) {
    if (W <= 0) return 0;
    const int TX = fmt == kF32 ? 128 : 256;
    return (long)((W + TX - 1) / TX) * (N / W) * (fmt == kF32 ? 48 : 24);
}
// Uup: band 13 of the scaled operator (16-bit: bands at k * pitch of the
// 25-band array, so Uup = A16 + 13 pitch; FP32: the 12 upper bands at m * pitch).
// XM: kXPlain (x updated, none pending), kXDefer (modes 4, 5) or kXFlush (the
// launch before deferred: modes 0 and 3 after a mode 4, and mode 2), so that
// each form carries only its own loads.
enum { kXPlain = 0, kXDefer = 1, kXFlush = 2 };
template <int FMT, int XM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_cgfused_symw(
    const void* __restrict__ Uup, size_t pitch, int W, int H, int N, const float* __restrict__ svec,
    float* __restrict__ x, float* r0, float* r1, float* p0, float* p1, float* Ap0, float* Ap1,
    const unsigned int* __restrict__ uedge, Scalars* sc, double* __restrict__ dpf, int seq, int mode, int verify) {
    constexpr int TY = 4;
    using SG = SymGeom<TY, FMT>;
    constexpr int R = SG::R, TPR = SG::TPR, TX = SG::TX, ST = SG::ST, SW = SG::SW, XR = SG::XR;
    // U ring of two tiles' rows plus the 2 above: the own rows of tile u + 1
    // arrive by direct-to-LDS loads while tile u's product reads its window
    constexpr int UR = 2 * TY + 2;
    constexpr int EPW = SG::EPW, EB = FMT == kF32 ? 4 : 2;  // band entries per word, bytes per entry
    static_assert(SW >= 2 * 64 + 2, "a staged U row holds the 128 words of two direct loads");
    constexpr int SPB = 2 / EPW;                             // side words per band and side
    using VR = std::conditional_t<R == 4, f32x4, f32x2>;
    auto ldR = [](__amdgpu_buffer_rsrc_t rs, long long i, bool skip = false) -> VR {
        if constexpr (R == 4) return buf_ld4(rs, i, skip);
        else return buf_ld2(rs, i, skip);
    };
    auto stR = [](__amdgpu_buffer_rsrc_t rs, long long i, VR v, bool skip = false) {
        if constexpr (R == 4) buf_st4(rs, i, v, skip);
        else buf_st2(rs, i, v, skip);
    };
    auto mkR = [](const float (&a)[R]) {
        VR v;
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = a[q];
        return v;
    };
    // p values of R consecutive window cells into the x ring (8-byte stores)
    auto put_p = [](float* d, const float (&pn)[R]) {
#pragma unroll
        for (int h = 0; h < R / 2; ++h) *reinterpret_cast<f32x2*>(d + 2 * h) = f32x2{pn[2 * h], pn[2 * h + 1]};
    };
    static_assert(TPR == 64 && XR == 2 * TY, "one wave per tile row; the x ring holds two tiles of rows");
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (sc->done != kRun || sc->verify) return;
    const bool first = mode == 1, timing = mode == 2 || mode == 5;
    FHand hin;
    if (first) {
        hin.rr_prev = sc->rz_prev;
        hin.beta = sc->beta;
        hin.iters = sc->iters;
        hin.npart = 0;
        hin.xpend = 0;
        hin.xa = 0.f;
    } else if (timing) {
        hin.rr_prev = 1.0;
        hin.beta = 0.5f;
        hin.iters = 0;
        hin.npart = 0;
        hin.xpend = XM == kXFlush;
        hin.xa = 0.5f;
    } else {
        hin = sc->fh[seq & 1];
    }
    // this launch defers its update of x; the launch before did (the host
    // alternates them, never two deferring launches in a row)
    constexpr bool defer = XM == kXDefer, xpend = XM == kXFlush;
    const int j = hin.iters;
    const int k = first ? j : j + 1;  // the step this launch makes
    const float* rin = (j & 1) ? r1 : r0;
    float* rout = (k & 1) ? r1 : r0;
    const float* Apin = (j & 1) ? Ap1 : Ap0;
    float* Apout = (k & 1) ? Ap1 : Ap0;
    const float* pold = (k & 1) ? p1 : p0;
    float* pnew = (k & 1) ? p0 : p1;
    __shared__ __attribute__((aligned(16))) float lds[XR * ST];
    __shared__ __attribute__((aligned(16))) unsigned int ub[12][UR][SW];
    // tile rows 0, 1: r_{j+1} and p_j of the next tile's output cells, held
    // from the tile that loaded them (each thread its own slot)
    __shared__ VR held[2][2 * TPR];
    const int ncol = (W + TX - 1) / TX, nrow = (H + TY - 1) / TY;
    int ubeg, uend;  // the run of column-major tiles of this workgroup (as k_cgstep_symw)
    {
        const int g = gridDim.x / kNumXcd, xcd = blockIdx.x % kNumXcd, l = blockIdx.x / kNumXcd;
        if (gridDim.x % kNumXcd == 0 && g >= ncol) {
            const int sg = g / ncol, nseg = kNumXcd * sg;
            if (l < ncol * sg) {
                const int col = l % ncol, seg = xcd * sg + l / ncol;
                ubeg = col * nrow + (int)((long long)seg * nrow / nseg);
                uend = col * nrow + (int)((long long)(seg + 1) * nrow / nseg);
            } else {
                ubeg = uend = 0;
            }
        } else {
            const long long T = (long long)ncol * nrow;
            ubeg = (int)(T * blockIdx.x / gridDim.x);
            uend = (int)(T * (blockIdx.x + 1) / gridDim.x);
        }
    }
    const int tx = (int)threadIdx.x % TPR, ty = (int)threadIdx.x / TPR;
    // this thread's window row among the 4 a tile loads: 4 + lrow (image row
    // y0 + 2 + lrow); tile rows 2, 3 load their own output rows (now), rows
    // 0, 1 the next tile's
    const int lrow = (ty + 2) & 3;
    const bool now = ty >= 2;
    constexpr int NS = 24 * SPB * TY;  // side words of the own U rows (row, band, side, word)
    constexpr int SE = (NS + kBlock - 1) / kBlock;
    struct Loads {
        unsigned int sv[SE];
        VR xr, xa, xp;     // r_j, Ap_j, p_j of the thread's R interior window cells
        float cr, ca, cp;  // ... of one side cell (threads 0 .. 15)
        VR xo, so;         // x and s of the thread's output cells
        VR xq;             // p_{j-1} of the output cells (kXFlush)
    };
    // every load and store goes through a range-checked buffer descriptor
    // (out-of-range: loads 0, stores dropped), so every path issues the same
    // memory instructions and the wait for a tile's loads never has to wait
    // for the stores of the tile before
    const __amdgpu_buffer_rsrc_t brr = buf_rsrc(rin, N), bra = buf_rsrc(Apin, N), brp = buf_rsrc(pold, N),
                                 brx = buf_rsrc(x, N), brs = buf_rsrc(svec, N), bwr = buf_rsrc(rout, N),
                                 bwp = buf_rsrc(pnew, N), bwa = buf_rsrc(Apout, N);
    // one range-checked descriptor over the 12 upper bands (24 pitch bytes <
    // 2^31): the band loads share one offset register, the band is the
    // scalar offset
    const __amdgpu_buffer_rsrc_t bu =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(Uup), 0, (int)(12 * EB * pitch), 0x00020000);
    const __amdgpu_buffer_rsrc_t be = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<unsigned int*>(uedge), 0, (int)(4LL * ((W + TX - 1) / TX) * H * 24 * SPB), 0x00020000);
    auto origin = [&](int u, int& x0, int& y0) {
        x0 = (u / nrow) * TX;
        y0 = (u % nrow) * TY;
    };
    // side cell e (< 16) of window rows w0 .. w0 + 3: row w0 + e / 4, column
    // 0, 1, TX + 2 or TX + 3
    auto side_gi = [&](int x0, int y0, int w0) {
        const int e = (int)threadIdx.x, c = e & 3;
        return (long long)(y0 - 2 + w0 + (e >> 2)) * W + (x0 - 2 + (c < 2 ? c : TX + c));
    };
    auto side_lx = [&]() {
        const int c = (int)threadIdx.x & 3;
        return c < 2 ? c : TX + c;
    };
    // the loads of tile u (none: the same instructions, all out of range, so
    // that every pass of the loop issues as many)
    // The own U rows of tile u straight into the ring (window row 2 + ty;
    // ut: the ring row of the tile's window row 0): two 4-byte loads per lane
    // and band, lane l taking interior words l and 64 + l of the row's 128
    // (entries x0 + EPW l ..; the interior starts at word 2 / EPW, after the
    // left side words), out of range -> 0.  None: the same loads, all out of
    // range (their zeros land in rows no tile reads).
    auto issue_u = [&](int u, bool none, int ut) {
        int x0, y0;
        origin(u, x0, y0);
        const int cy = y0 + ty;
        {
            const int wr = ut + 2 + ty < UR ? ut + 2 + ty : ut + 2 + ty - UR;
            const int lane = (int)threadIdx.x & 63;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int xw = x0 + EPW * (64 * h + lane);
                const unsigned off =
                    (xw < W && cy < H && !none) ? 4u * (unsigned)(((size_t)cy * W + xw) / EPW) : 0x80000000u;
#pragma unroll
                for (int m = 0; m < 12; ++m)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        bu, (__attribute__((address_space(3))) void*)&ub[m][wr][2 / EPW + 64 * h], 4, (int)off,
                        (int)(EB * m * pitch), 0, kNonTemporal);
            }
        }
    };
    // the rest of tile u's loads, into registers
    auto issue = [&](int u, Loads& L, bool none) {
        int x0, y0;
        origin(u, x0, y0);
        const int cx = x0 + R * tx, cy = y0 + ty;
        const bool valid = cx < W && cy < H && !none;
        const size_t ib = (size_t)cy * W + cx;
        // side words of the own rows: 96 contiguous words of the edge array
        // (rows of the image only; a row below it feeds no output cell)
#pragma unroll
        for (int q = 0; q < SE; ++q) {
            const int e = (int)threadIdx.x + q * kBlock;
            const bool in = e < NS && y0 + e / (24 * SPB) < H && !none;
            const unsigned off =
                in ? 4u * ((unsigned)((x0 / TX) * H + y0) * (24u * SPB) + (unsigned)e) : 0x80000000u;
            L.sv[q] = __builtin_amdgcn_raw_buffer_load_b32(be, (int)off, 0, 0);
        }
        if constexpr (!defer) L.xo = ldR(brx, (long long)ib, !valid || first);
        L.so = ldR(brs, (long long)ib, !valid);
        if constexpr (xpend) L.xq = ldR(bwp, (long long)ib, !valid);
        const long long gi = (long long)(y0 + 2 + lrow) * W + cx;
        L.xr = ldR(brr, gi, none);
        L.xa = ldR(bra, gi, first || none);
        L.xp = ldR(brp, gi, none);
        const long long gs = side_gi(x0, y0, 4);
        const bool sk = threadIdx.x >= 16 || none;
        L.cr = buf_ld(brr, gs, sk);
        L.ca = buf_ld(bra, gs, sk || first);
        L.cp = buf_ld(brp, gs, sk);
    };
    // software pipeline: the loads of tile u + 1 are issued while tile u is
    // computed (the first tile's while the previous launch's sums are folded)
    // The rows a run's first tile shares with the tile above (U window rows
    // 0, 1; x window rows 0 .. 3, one row per thread as the tile above would
    // have loaded them).  The first tile's are loaded with its other loads, so
    // they are in flight while the previous launch's sums are folded.
    constexpr int PE = (2 * SW + kBlock - 1) / kBlock;
    struct Pre {
        unsigned int u[PE][12];
        VR a, b, c;
        float sr, sa, sp;
    };
    auto load_pre = [&](int u, Pre& P) {
        int x0, y0;
        origin(u, x0, y0);
#pragma unroll
        for (int k = 0; k < PE; ++k) {
            const int e = (int)threadIdx.x + k * kBlock;
            const int wy = e / SW, ww = e - wy * SW;
            const long long gi = (long long)(y0 - 2 + wy) * W + (x0 - 2 + EPW * ww);
            const bool in = e < 2 * SW && gi >= 0 && gi < N;
#pragma unroll
            for (int m = 0; m < 12; ++m) {
                const unsigned off =
                    in ? (unsigned)EB * ((unsigned)m * (unsigned)pitch + (unsigned)gi) : 0x80000000u;
                P.u[k][m] = __builtin_amdgcn_raw_buffer_load_b32(bu, (int)off, 0, 0);
            }
        }
        const long long gi = (long long)(y0 - 2 + lrow) * W + (x0 + R * tx);
        P.a = ldR(brr, gi);
        P.b = ldR(bra, gi, first);
        P.c = ldR(brp, gi);
        const long long gs = side_gi(x0, y0, 0);
        const bool sk = threadIdx.x >= 16;
        P.sr = buf_ld(brr, gs, sk);
        P.sa = buf_ld(bra, gs, sk || first);
        P.sp = buf_ld(brp, gs, sk);
    };
    Loads L;
    Pre P;
    if (ubeg < uend && mode != 3) {
        issue_u(ubeg, false, 0);
        issue(ubeg, L, false);
        load_pre(ubeg, P);
    }
    float alpha = 0.f, beta = hin.beta;
    FHand hout{hin.rr_prev, hin.beta, j, (int)gridDim.x, 0, 0.f};
    if (timing) {
        alpha = 0.5f;
    } else if (!first) {
        double t[5];
        fold_all<5>(dpf + (size_t)(seq & 1) * 5 * kMaxGrid, hin.npart, t);
        const double pAp = t[0], rr = t[1], rAp = t[2], ApAp = t[3], rru = t[4];
        int stop = kRun, ver = 0;
        if (!isfinite(pAp) || !isfinite(rr) || !isfinite(rAp) || !isfinite(ApAp) || !isfinite(rru)) {
            stop = kNonFinite;
        } else if (sc->tol2 >= 0 && rru <= sc->tol2) {
            if (verify) ver = 1;
            else stop = kConverged;
        } else if (j >= sc->maxit) {
            stop = kMaxIt;
        } else if (!(pAp > 0)) {
            stop = kBreakdown;
        }
        if (stop != kRun || ver) {
            if constexpr (xpend) {
                // the update deferred by the launch before (p_{j-1} is in the
                // buffer this launch would have written), so that x matches r_j
                const int n4 = N >> 2;
                for (int i = blockIdx.x * kBlock + (int)threadIdx.x; i < n4; i += gridDim.x * kBlock) {
                    const f32x4 sv4 = reinterpret_cast<const f32x4*>(svec)[i];
                    const f32x4 qv = reinterpret_cast<const f32x4*>(pnew)[i];
                    f32x4 xv = reinterpret_cast<const f32x4*>(x)[i];
#pragma unroll
                    for (int q = 0; q < 4; ++q) xv[q] = fmaf(hin.xa * sv4[q], qv[q], xv[q]);
                    reinterpret_cast<f32x4*>(x)[i] = xv;
                }
            }
            if (lead) {
                sc->rz = rr;
                sc->rz_prev = hin.rr_prev;
                sc->rr = rru;
                sc->beta = hin.beta;
                sc->pAp = pAp;
                sc->iters = j;
                sc->pend = 0;
                sc->verify = ver;
                sc->done = stop;
            }
            return;
        }
        alpha = (float)(rr / pAp);
        const double a = (double)alpha, rr1 = rr + 2.0 * a * rAp + a * a * ApAp;
        beta = rr1 > 0 ? (float)(rr1 / rr) : 0.f;
        hout.rr_prev = rr;
        hout.beta = beta;
        hout.iters = j + 1;
    }
    if constexpr (defer) {
        hout.xpend = 1;
        hout.xa = alpha;
    }
    if (mode == 3) {
        // own cells only, float4 groups over the whole vector (and the update
        // of x deferred by the launch before, if any)
        double w[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const int n4 = N >> 2;
        for (int i = blockIdx.x * kBlock + (int)threadIdx.x; i < n4; i += gridDim.x * kBlock) {
            f32x4 xv = reinterpret_cast<const f32x4*>(x)[i];
            const f32x4 sv4 = reinterpret_cast<const f32x4*>(svec)[i];
            const f32x4 pv = reinterpret_cast<const f32x4*>(pold)[i], rv = reinterpret_cast<const f32x4*>(rin)[i];
            const f32x4 av = reinterpret_cast<const f32x4*>(Apin)[i];
            if constexpr (xpend) {
                const f32x4 qv = reinterpret_cast<const f32x4*>(pnew)[i];
#pragma unroll
                for (int q = 0; q < 4; ++q) xv[q] = fmaf(hin.xa * sv4[q], qv[q], xv[q]);
            }
            f32x4 xn, rn;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                xn[q] = fmaf(alpha * sv4[q], pv[q], xv[q]);
                rn[q] = fmaf(alpha, av[q], rv[q]);
                const double rd = rn[q];
                const float un = rn[q] / sv4[q];
                w[1] += rd * rd;
                w[4] += (double)un * (double)un;
            }
            reinterpret_cast<f32x4*>(x)[i] = xn;
            reinterpret_cast<f32x4*>(rout)[i] = rn;
        }
        store_partials<5>(w, dpf + (size_t)((seq + 1) & 1) * 5 * kMaxGrid);
        if (lead) sc->fh[(seq + 1) & 1] = hout;
        return;
    }
    // r_{j+1} and p_{j+1} of a loaded cell
    auto advance = [&](float rv, float av, float pv, float& rn) {
        rn = first ? rv : fmaf(alpha, av, rv);
        return beta != 0.f ? fmaf(beta, pv, -rn) : -rn;
    };
    int xtop = 0, utop = 0;  // LDS rows of window row 0 (x ring of XR rows, U ring of UR rows)
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // The first tile's loads are complete on every path into the loop, so the
    // compiler's wait for a tile's loads in the loop only has to cover the
    // stores issued after them (vmcnt(0) alone: expcnt and lgkmcnt at their
    // maxima).
    __builtin_amdgcn_s_waitcnt(0xF70);
    for (int u = ubeg; u < uend; ++u) {
        int x0, y0;
        origin(u, x0, y0);
        const bool start = u == ubeg || y0 == 0;
        __syncthreads();  // the previous tile is done with lds and ub
        // the U ring advances by a tile's rows on every tile (its own rows were
        // loaded there during the tile before); the x ring restarts at a run's
        // or column's first tile
        if (u != ubeg) utop = utop + TY < UR ? utop + TY : utop + TY - UR;
        // the side words of the own U rows (window rows 2 .. TY + 1; their
        // interior words are already in the ring)
        {
#pragma unroll
            for (int q = 0; q < SE; ++q) {
                const int e = (int)threadIdx.x + q * kBlock;
                if (e < NS) {
                    // (row, band, side, word): words 0 .. SPB - 1 (left) or SW - SPB .. SW - 1 (right)
                    const int row = e / (24 * SPB), rem = e % (24 * SPB);
                    const int m = rem / (2 * SPB), side = (rem / SPB) & 1, wd = rem % SPB;
                    const int wy = utop + 2 + row, ws = wy < UR ? wy : wy - UR;
                    ub[m][ws][side ? SW - SPB + wd : wd] = L.sv[q];
                }
            }
        }
        // the next tile's own U rows, once this tile's side words are written:
        // ring rows this tile does not read, and which the tile before read
        // only up to the barrier passed above (the ring accesses after them --
        // a run's first rows of a new column, the barrier before the product
        // -- wait for them)
        issue_u(u + 1, u + 1 >= uend, utop + TY < UR ? utop + TY : utop + TY - UR);
        if (start) {
            // the rows shared with the tile above: U window rows 0, 1 and x
            // window rows 0 .. 3 (as the tile above would have loaded them)
            xtop = 0;
            if (u != ubeg) load_pre(u, P);  // a run that enters the next column
#pragma unroll
            for (int k = 0; k < PE; ++k) {
                const int e = (int)threadIdx.x + k * kBlock;
                if (e < 2 * SW) {
                    const int wy = e / SW, ww = e - wy * SW;
                    const int ws = utop + wy < UR ? utop + wy : utop + wy - UR;
#pragma unroll
                    for (int m = 0; m < 12; ++m) ub[m][ws][ww] = P.u[k][m];
                }
            }
            {
                const VR a = P.a, b = P.b, c = P.c;
                float rn[R], pn[R];
#pragma unroll
                for (int q = 0; q < R; ++q) pn[q] = advance(a[q], b[q], c[q], rn[q]);
                put_p(lds + lrow * ST + 2 + R * tx, pn);
                if (!now) {
                    held[0][threadIdx.x] = mkR(rn);
                    held[1][threadIdx.x] = c;
                }
            }
            if (threadIdx.x < 16) {
                float rn;
                const float pn = advance(P.sr, P.sa, P.sp, rn);
                lds[((int)threadIdx.x >> 2) * ST + side_lx()] = pn;
            }
        } else {
            // window rows TY.. of the tile above are rows 0.. of this one
            xtop = (xtop + TY) & (XR - 1);
        }
        // x window rows 4 .. 7: p_{j+1}.  The update of the output cells (x,
        // r_{j+1} and their squared norms) is done here, so that only r_{j+1}
        // stays live through the product.
        const int cx = x0 + R * tx, cy = y0 + ty;
        const bool valid = cx < W && cy < H;
        const int ib = cy * W + cx;
        float cur_r[R];
        {
            float rn[R], pn[R], cur_p[R];
#pragma unroll
            for (int q = 0; q < R; ++q) pn[q] = advance(L.xr[q], L.xa[q], L.xp[q], rn[q]);
            put_p(lds + ((xtop + 4 + lrow) & (XR - 1)) * ST + 2 + R * tx, pn);
            if (now) {
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    cur_r[q] = rn[q];
                    cur_p[q] = L.xp[q];
                }
            } else {
                const VR hr = held[0][threadIdx.x], hp = held[1][threadIdx.x];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    cur_r[q] = hr[q];
                    cur_p[q] = hp[q];
                }
                held[0][threadIdx.x] = mkR(rn);
                held[1][threadIdx.x] = L.xp;
            }
            if (threadIdx.x < 16) {
                float rs;
                lds[((xtop + 4 + ((int)threadIdx.x >> 2)) & (XR - 1)) * ST + side_lx()] =
                    advance(L.cr, L.ca, L.cp, rs);
            }
            {
                // x_{j+1} = x_j + alpha_j s p_j, after x_j = x_{j-1} + alpha_{j-1} s p_{j-1}
                // where the launch before deferred that
                if constexpr (!defer) {
                    float xn[R];
#pragma unroll
                    for (int q = 0; q < R; ++q) {
                        float xb = L.xo[q];
                        if constexpr (xpend) xb = fmaf(hin.xa * L.so[q], L.xq[q], xb);
                        xn[q] = fmaf(alpha * L.so[q], cur_p[q], xb);
                    }
                    stR(brx, ib, mkR(xn), !valid || first);
                }
                stR(bwr, ib, mkR(cur_r), !valid || first);
            }
            if (valid) {
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    if (ib + q < N) {
                        const double rd = cur_r[q];
                        const float un = cur_r[q] / L.so[q];
                        v[1] += rd * rd;
                        v[4] += (double)un * (double)un;
                    }
                }
            }
        }
        __syncthreads();
        issue(u + 1, L, u + 1 >= uend);
        float yv[R] = {}, pc[R] = {};
        if (valid) {
            int xrow[5], urow[3];
#pragma unroll
            for (int q = 0; q < 5; ++q) xrow[q] = ((xtop + ty + q) & (XR - 1)) * ST + R * tx;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int lr = utop + ty + 2 - d;
                urow[d] = lr < UR ? lr : lr - UR;
            }
            sym_tile<TY, FMT, UR>(lds, xrow, ub, urow, tx, yv, pc);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (ib + q < N) {
                    const double yd = yv[q];
                    v[0] += (double)(pc[q] * yv[q]);
                    v[2] += (double)cur_r[q] * yd;
                    v[3] += yd * yd;
                }
            }
        }
        stR(bwp, ib, mkR(pc), !valid);
        stR(bwa, ib, mkR(yv), !valid);
    }
    store_partials<5>(v, dpf + (size_t)((seq + 1) & 1) * 5 * kMaxGrid);
    if (lead && !timing) sc->fh[(seq + 1) & 1] = hout;
}

// Publish the fused iteration's live state (slot seq & 1) into the Scalars
// and run the stopping test (one workgroup; before the host reads the
// Scalars or replaces the residual).
__global__ __launch_bounds__(kBlock) void k_fused_fold(Scalars* sc, const double* __restrict__ dpf, int seq,
                                                       int verify) {
    if (sc->done != kRun || sc->verify) return;
    const FHand h = sc->fh[seq & 1];
    double t[5];
    fold_all<5>(dpf + (size_t)(seq & 1) * 5 * kMaxGrid, h.npart, t);
    if (threadIdx.x != 0) return;
    sc->rz = t[1];
    sc->rz_prev = h.rr_prev;
    sc->beta = h.beta;
    sc->pAp = t[0];
    sc->iters = h.iters;
    sc->pend = 0;
    sc->rr = t[4];
    // x lags r after a launch that deferred its update (the host folds only
    // after one that did not): no stop here, the next launch decides it and
    // brings x up to date
    if (h.xpend) return;
    if (!isfinite(t[0]) || !isfinite(t[1]) || !isfinite(t[4])) sc->done = kNonFinite;
    else if (sc->tol2 >= 0 && t[4] <= sc->tol2) {
        if (verify) sc->verify = 1;
        else sc->done = kConverged;
    } else if (h.iters >= sc->maxit) {
        sc->done = kMaxIt;
    }
}

// K2: x += alpha p (scaled: alpha s p), r += alpha Ap, <r,z>, ||r||^2 (scaled:
// the unscaled residual (r/s)^2).  VEC: every pointer is 16-byte aligned and
// the body moves float4s (the kernel is a pure stream: wide loads keep enough
// bytes in flight to reach HBM bandwidth).
template <int PM>
__device__ __forceinline__ void upd_elem(float& x, float& r, float d, float p, float ap, float alpha, double& v0,
                                         double& v1) {
    if (PM == kScaled) x = fmaf(alpha * d, p, x);
    else x = fmaf(alpha, p, x);
    const float rn = fmaf(alpha, ap, r);
    r = rn;
    if (PM == kJacobi) {
        v0 += (double)(rn * (d * rn));
        v1 += (double)(rn * rn);
    } else if (PM == kScaled) {
        const float u = rn / d;  // unscaled residual
        v0 += (double)(rn * rn);
        v1 += (double)(u * u);
    } else {
        v0 += (double)(rn * rn);
    }
}

template <int PM, bool VEC>
__global__ __launch_bounds__(kBlock) void k_cgupdate(float* __restrict__ x, float* __restrict__ r,
                                                     const float* __restrict__ dinv, const float* __restrict__ p0,
                                                     const float* __restrict__ p1, const float* __restrict__ Ap,
                                                     int N, Scalars* sc, const double* __restrict__ dpa,
                                                     double* __restrict__ dpb) {
    const Hand h = sc->hand;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (h.mode != kHandRun) {
        if (h.mode == kHandStop && lead) publish(sc, h, h.done, h.verify, h.iters, 0, 0);
        return;
    }
    double t[1];
    fold_all<1>(dpa, h.npart, t);
    const double pAp = t[0];
    if (!isfinite(pAp) || !(pAp > 0)) {
        if (lead) {
            publish(sc, h, isfinite(pAp) ? kBreakdown : kNonFinite, 0, h.iters, 0, 0);
            sc->pAp = pAp;
        }
        return;
    }
    const float alpha = (float)(h.rz / pAp);
    const float* __restrict__ p = (h.iters & 1) ? p0 : p1;  // the p_new K1 wrote
    double v[2] = {0.0, 0.0};
    const int tid = blockIdx.x * kBlock + threadIdx.x, nth = gridDim.x * kBlock;
    if (VEC) {
        const int n4 = N >> 2;
        for (int i = tid; i < n4; i += nth) {
            f32x4 xv = reinterpret_cast<const f32x4*>(x)[i];
            f32x4 rv = reinterpret_cast<const f32x4*>(r)[i];
            const f32x4 pv = reinterpret_cast<const f32x4*>(p)[i];
            const f32x4 av = reinterpret_cast<const f32x4*>(Ap)[i];
            f32x4 dv = {1.f, 1.f, 1.f, 1.f};
            if (PM != kPlain) dv = reinterpret_cast<const f32x4*>(dinv)[i];
            float xs[4] = {xv.x, xv.y, xv.z, xv.w}, rs[4] = {rv.x, rv.y, rv.z, rv.w};
            const float ds[4] = {dv.x, dv.y, dv.z, dv.w}, ps[4] = {pv.x, pv.y, pv.z, pv.w};
            const float as[4] = {av.x, av.y, av.z, av.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) upd_elem<PM>(xs[q], rs[q], ds[q], ps[q], as[q], alpha, v[0], v[1]);
            xv = f32x4{xs[0], xs[1], xs[2], xs[3]};
            rv = f32x4{rs[0], rs[1], rs[2], rs[3]};
            reinterpret_cast<f32x4*>(x)[i] = xv;
            reinterpret_cast<f32x4*>(r)[i] = rv;
        }
        for (int i = (n4 << 2) + tid; i < N; i += nth) {
            float xi = x[i], ri = r[i];
            upd_elem<PM>(xi, ri, PM != kPlain ? dinv[i] : 1.f, p[i], Ap[i], alpha, v[0], v[1]);
            x[i] = xi;
            r[i] = ri;
        }
    } else {
        for (int i = tid; i < N; i += nth) {
            float xi = x[i], ri = r[i];
            upd_elem<PM>(xi, ri, PM != kPlain ? dinv[i] : 1.f, p[i], Ap[i], alpha, v[0], v[1]);
            x[i] = xi;
            r[i] = ri;
        }
    }
    if (PM == kPlain) v[1] = v[0];
    store_partials<2>(v, dpb);
    if (lead) {
        publish(sc, h, kRun, 0, h.iters + 1, 1, (int)gridDim.x);
        sc->pAp = pAp;
        sc->alpha = alpha;
    }
}

// Fold pending update partials into the Scalars (one workgroup; the host
// launches it before reading the final state).
__global__ __launch_bounds__(kBlock) void k_fold(Scalars* sc, const double* __restrict__ dpb, int verify) {
    if (sc->done != kRun || sc->verify || !sc->pend) return;
    double t[2];
    fold_all<2>(dpb, sc->npart, t);
    if (threadIdx.x != 0) return;
    sc->pend = 0;
    if (!isfinite(t[0]) || !isfinite(t[1])) {
        sc->done = kNonFinite;
        return;
    }
    sc->rz_prev = sc->rz;
    sc->beta = (float)(t[0] / sc->rz);
    sc->rz = t[0];
    sc->rr = t[1];
    if (sc->tol2 >= 0 && sc->rr <= sc->tol2) {
        if (verify) sc->verify = 1;
        else sc->done = kConverged;
    } else if (sc->iters >= sc->maxit) {
        sc->done = kMaxIt;
    }
}

// FP64-accumulated true residual t = A x - b.  op 0: measure only (g0 = ||t||^2);
// op 1: replace (r = t, or s*t for the scaled system; <r,z>, ||t||^2);
// op 2: start the scaled system (as 1 with beta = 0).  XT: float for the
// solver's own x, double for the FP64 master x of refine mode 2.
template <int R, int PM, class XT>
__global__ __launch_bounds__(kBlock) void k_trueres25(const float* __restrict__ A, size_t pitch, int W, int N,
                                                      const XT* __restrict__ x, const float* __restrict__ b,
                                                      const float* __restrict__ dinv, float* __restrict__ r, int op,
                                                      Scalars* sc, double* partials, unsigned int* counter) {
    if (op != 0 && sc->done != kRun) return;
    __shared__ __attribute__((aligned(16))) XT lds[Tile<R>::LDS_FLOATS];
    const int ntiles = (N + Tile<R>::ROWS - 1) / Tile<R>::ROWS;
    const TileRange tr = xcd_tiles(ntiles);
    constexpr int SEG = Tile<R>::SEG;
    double v[3] = {0.0, 0.0, 0.0};
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int i0 = t * Tile<R>::ROWS;
        __syncthreads();
        int nz = 0;
        stage5<R>(lds, i0, W, N, [&](int g) {
            const XT v = x[g];
            nz |= v != XT(0);
            return v;
        });
        // a window of zeros (a zero initial guess) needs no band loads: A x = 0
        nz = __syncthreads_or(nz);
        const int ib = i0 + R * threadIdx.x;
        if (ib < N) {
            const XT* xs = lds + R * threadIdx.x;
            double acc[R];
#pragma unroll
            for (int j = 0; j < R; ++j) acc[j] = 0.0;
            if (nz) {
#pragma unroll
                for (int k = 0; k < 25; ++k) {
                    Raw<R, kF32> a;
                    load_band<R, kF32>(A, (size_t)k * pitch + (size_t)ib, a);
                    const int s = k / 5, dx = k % 5;
#pragma unroll
                    for (int j = 0; j < R; ++j) acc[j] = fma((double)a.get(j), (double)xs[s * SEG + j + dx], acc[j]);
                }
            }
            float bb[R], dd[R], out[R], ro[R];
            load_rows<R>(b, ib, N, bb);
            if (PM != kPlain) load_rows<R>(dinv, ib, N, dd);
            if (op == 1) load_rows<R>(r, ib, N, ro);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const double tj = acc[j] - (double)bb[j];
                float rj = (float)tj;
                if (PM == kScaled) rj = (float)(tj * (double)dd[j]);
                out[j] = rj;
                if (ib + j < N) {
                    v[1] += tj * tj;
                    if (PM == kJacobi) v[0] += (double)(rj * (dd[j] * rj));
                    else v[0] += (double)(rj * rj);
                    if (op == 1) v[2] += (double)(rj - ro[j]) * (double)(rj - ro[j]);
                }
            }
            if (op != 0) store_rows<R>(r, ib, N, out);
        }
    }
    grid_reduce<3>(v, partials, counter, [&](double* tt) {
        sc->g0 = tt[1];
        if (op == 0) return;
        fin_residual(sc, tt[0], tt[1], op == 2, true, tt[2], PM == kScaled ? tt[0] : tt[1]);
    });
}

// k_trueres25 on 2-D pixel tiles (Geom2D: the FP32 step's staging, which
// keeps the x halo in L2); the same sums in the same order.  As in the FP32
// step kernel, a tile's 25 band loads are in flight while its x window is
// staged, and each workgroup stores its partial sums for k_trueres_fin (no
// fenced reduction, so the grid can hold 3 workgroups per CU).
// r_odd: the fused iteration's second residual buffer, which holds the
// residual after an odd number of iterations (else null).  xzero: the caller
// knows that k_bnorm saw x = 0 (the first residual of a solve): r = -b, no
// band loads.
template <class G, int PM, class XT>
__global__ __launch_bounds__(kBlock) void k_trueres2d(const float* __restrict__ A, size_t pitch, int W, int H, int N,
                                                      const XT* __restrict__ x, const float* __restrict__ b,
                                                      const float* __restrict__ dinv, float* __restrict__ r,
                                                      float* __restrict__ r_odd, int op, int xzero, Scalars* sc,
                                                      double* __restrict__ partials) {
    if (op != 0 && sc->done != kRun) return;
    if (r_odd && (sc->iters & 1)) r = r_odd;
    const bool zero = xzero && sc->xzero;
    constexpr int R = G::R, ST = G::STRIDE;
    __shared__ __attribute__((aligned(16))) XT lds[G::LDS_FLOATS];
    const TileRange tr = xcd_tiles(G::ntiles(W, H, N));
    double v[3] = {0.0, 0.0, 0.0};
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        G g;
        g.at(t, W, H, N);
        Bands<R, kF32> av;
        if (g.valid && !zero) load_bands<R, kF32, kDiagStored>(A, pitch, g.ib, av);
        __syncthreads();  // the previous tile is done with lds
        if (!zero) g.stage(lds, W, H, N, [&](int i) { return x[i]; });
        __syncthreads();
        if (!g.valid) continue;
        double acc[R];
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = 0.0;
        if (!zero) {
            const XT* xs = g.xs(lds);
#pragma unroll
            for (int k = 0; k < 25; ++k) {
                const int s = k / 5, dx = k % 5;
#pragma unroll
                for (int j = 0; j < R; ++j)
                    acc[j] = fma((double)av.v[k].get(j), (double)xs[s * ST + j + dx], acc[j]);
            }
        }
        float bb[R], dd[R], out[R], ro[R];
        load_rows<R>(b, g.ib, N, bb);
        if (PM != kPlain) load_rows<R>(dinv, g.ib, N, dd);
        if (op == 1) load_rows<R>(r, g.ib, N, ro);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const double tj = acc[j] - (double)bb[j];
            float rj = (float)tj;
            if (PM == kScaled) rj = (float)(tj * (double)dd[j]);
            out[j] = rj;
            if (g.ib + j < N) {
                v[1] += tj * tj;
                if (PM == kJacobi) v[0] += (double)(rj * (dd[j] * rj));
                else v[0] += (double)(rj * rj);
                if (op == 1) v[2] += (double)(rj - ro[j]) * (double)(rj - ro[j]);
            }
        }
        if (op != 0) store_rows<R>(r, g.ib, N, out);
    }
    store_partials<3>(v, partials);
}

// k_trueres2d on long row runs (as the banded SpMV k_spmv_seg): tiles of
// 256 G x 4 pixels, one wave per tile row, each lane G runs of 4 pixels 1 KB
// apart, the bands in five passes of five (window row dy, then dx), each
// pass's loads issued before its FP64 FMAs.  Every row's sum is the same FMA
// chain in band order k = 0..24 as k_trueres2d, and so are the residual and
// its roundings; only the partition of the partial sums differs.  The x
// window (XT: float, or the FP64 master) is staged in LDS with window column c
// at LDS column c.  W % 4 == 0, 16-byte aligned vectors, 25 band planes below
// 2^31 bytes.
template <int G, int PM, class XT>
__global__ __launch_bounds__(kBlock) void k_trueres_seg(const float* __restrict__ A, size_t pitch, int W, int H,
                                                        int N, const XT* __restrict__ x, const float* __restrict__ b,
                                                        const float* __restrict__ dinv, float* __restrict__ r,
                                                        float* __restrict__ r_odd, int op, int xzero, Scalars* sc,
                                                        double* __restrict__ partials) {
    if (op != 0 && sc->done != kRun) return;
    if (r_odd && (sc->iters & 1)) r = r_odd;
    const bool zero = xzero && sc->xzero;
    constexpr int TX = 256 * G, TY = 4, ST = TX + 4, C4 = TX / 4;
    constexpr int IE = (TY + 4) * C4 / kBlock;  // interior 4-cell groups staged per thread
    static_assert((TY + 4) * C4 % kBlock == 0, "whole staging rounds");
    __shared__ __attribute__((aligned(16))) XT lds[(TY + 4) * ST];
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int ncol = (W + TX - 1) / TX, nrow = (H + TY - 1) / TY;
    const TileRange tr = xcd_tiles(ncol * nrow);
    const __amdgpu_buffer_rsrc_t bx = __builtin_amdgcn_make_buffer_rsrc(const_cast<XT*>(x), 0,
                                                                        (int)(sizeof(XT) * (size_t)N), 0x00020000);
    const __amdgpu_buffer_rsrc_t bb = buf_rsrc(b, N), bd = buf_rsrc(dinv, PM != kPlain ? N : 0),
                                 br = buf_rsrc(r, N);
    const __amdgpu_buffer_rsrc_t ba =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(A), 0, (int)(25 * 4 * pitch), 0x00020000);
    // 4 consecutive x cells from element i (i % 4 == 0; out of range: 0)
    auto ldx4 = [&](long long i, XT (&v)[4]) {
        const unsigned off = (unsigned)i * (unsigned)sizeof(XT);
        if constexpr (sizeof(XT) == 4) {
            const f32x4 t = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(bx, (int)off, 0, 0));
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = t[q];
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const u32x4 t = __builtin_amdgcn_raw_buffer_load_b128(bx, (int)(off + 16u * h), 0, 0);
                v[2 * h] = __builtin_bit_cast(double, u32x2{t.x, t.y});
                v[2 * h + 1] = __builtin_bit_cast(double, u32x2{t.z, t.w});
            }
        }
    };
    auto ldx1 = [&](long long i) -> XT {
        const unsigned off = (unsigned)i * (unsigned)sizeof(XT);
        if constexpr (sizeof(XT) == 4) return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(bx, (int)off, 0, 0));
        else return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(bx, (int)off, 0, 0));
    };
    double v[3] = {0.0, 0.0, 0.0};
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int x0 = (t % ncol) * TX, y0 = (t / ncol) * TY, y = y0 + w;
        unsigned off[G];  // byte offset of each run in a band plane (out of range: invalid)
        bool val[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int xc = x0 + 256 * g + 4 * lane;
            val[g] = xc < W && y < H;
            off[g] = val[g] ? 4u * (unsigned)(y * W + xc) : 0x80000000u;
        }
        f32x4 a[5][G];
        auto issue = [&](int dy) {
#pragma unroll
            for (int dx = 0; dx < 5; ++dx)
#pragma unroll
                for (int g = 0; g < G; ++g)
                    a[dx][g] = __builtin_bit_cast(
                        f32x4, __builtin_amdgcn_raw_buffer_load_b128(ba, (int)(zero ? 0x80000000u : off[g]),
                                                                     (int)(4 * (5 * dy + dx) * pitch), kNonTemporal));
        };
        __syncthreads();  // the previous tile is done with lds
        if (!zero) {
            XT xi[IE][4];
#pragma unroll
            for (int q = 0; q < IE; ++q) {
                const int e = (int)threadIdx.x + q * kBlock, ly = e / C4, c4 = e - ly * C4;
                ldx4((long long)(y0 - 2 + ly) * W + x0 + 4 * c4, xi[q]);
            }
            XT xh = XT(0);
            const int hly = (int)threadIdx.x >> 2, hc = (int)threadIdx.x & 3, hlx = hc < 2 ? hc : TX + hc;
            if (threadIdx.x < 4 * (TY + 4)) xh = ldx1((long long)(y0 - 2 + hly) * W + x0 - 2 + hlx);
            issue(0);
#pragma unroll
            for (int q = 0; q < IE; ++q) {
                const int e = (int)threadIdx.x + q * kBlock, ly = e / C4, c4 = e - ly * C4;
#pragma unroll
                for (int c = 0; c < 4; ++c) lds[ly * ST + 2 + 4 * c4 + c] = xi[q][c];
            }
            if (threadIdx.x < 4 * (TY + 4)) lds[hly * ST + hlx] = xh;
        }
        __syncthreads();
        double acc[G][4];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[g][j] = 0.0;
        if (!zero) {
#pragma unroll
            for (int dy = 0; dy < 5; ++dy) {
                if (dy > 0) issue(dy);
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const XT* sx = lds + (w + dy) * ST + 256 * g + 4 * lane;
#pragma unroll
                    for (int dx = 0; dx < 5; ++dx)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[g][j] = fma((double)a[dx][g][j], (double)sx[j + dx], acc[g][j]);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const unsigned o = val[g] ? off[g] : 0x80000000u;
            const f32x4 bv = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(bb, (int)o, 0, 0));
            f32x4 dv = {0.f, 0.f, 0.f, 0.f}, rv = {0.f, 0.f, 0.f, 0.f};
            if (PM != kPlain) dv = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(bd, (int)o, 0, 0));
            if (op == 1) rv = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(br, (int)o, 0, 0));
            f32x4 out;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double tj = acc[g][j] - (double)bv[j];
                float rj = (float)tj;
                if (PM == kScaled) rj = (float)(tj * (double)dv[j]);
                out[j] = rj;
                if (val[g]) {
                    v[1] += tj * tj;
                    if (PM == kJacobi) v[0] += (double)(rj * (dv[j] * rj));
                    else v[0] += (double)(rj * rj);
                    if (op == 1) v[2] += (double)(rj - rv[j]) * (double)(rj - rv[j]);
                }
            }
            if (op != 0) __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, out), br, (int)o, 0, 0);
        }
    }
    store_partials<3>(v, partials);
}

// The reduction of k_trueres2d (one workgroup): g0, and for op 1 / 2 the
// replacement and stopping test of k_trueres25's finaliser.
__global__ __launch_bounds__(kBlock) void k_trueres_fin(Scalars* sc, const double* __restrict__ partials, int npart,
                                                        int op, int scaled) {
    if (op != 0 && sc->done != kRun) return;
    double t[3];
    fold_all<3>(partials, npart, t);
    if (threadIdx.x != 0) return;
    sc->g0 = t[1];
    if (op == 0) return;
    fin_residual(sc, t[0], t[1], op == 2, true, t[2], scaled ? t[0] : t[1]);
}

// Mixed-precision operator: As = S A S with S = diag(1/sqrt(a_ii)) (unit
// diagonal, band 12 not stored), rounded to FP16 or BF16, in two streaming
// passes.  k_svec: s (1 where a_ii <= 0).  k_scale16: each thread 4
// consecutive rows, every band's 16-byte chunk and the neighbour scales of
// each band issued before any conversion; the neighbour scales (s at i + o,
// 0 outside the vector, where the band entry is 0 too) come from L2, since s
// was just written.  (a_ik s_i) s_{i+o} in that order.
// SYM: only the upper bands 13..24 (the symmetric step kernel's form; W % 4
// == 0): the neighbours of rows ib .. ib + 3 in window row dy = 0, 1, 2 are
// cells ib + dy W - 2 .. ib + dy W + 5, read as three aligned 16-byte groups
// per window row through a range-checked buffer (a group outside the vector
// reads 0: N % 4 == 0, so a group is all inside or all outside).
template <bool SYM>
struct ScaleWin {
    f32x4 g[SYM ? 3 : 1][SYM ? 3 : 1];
};
// the 3 x 3 aligned groups of s around rows ib .. ib + 3 (window rows dy = 0..2)
__device__ __forceinline__ void load_swin(__amdgpu_buffer_rsrc_t bs, int W, int ib, ScaleWin<true>& sw) {
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int c = 0; c < 3; ++c) sw.g[dy][c] = buf_ld4(bs, (long long)ib + (long long)dy * W - 4 + 4 * c);
}
// s at row ib + j + dy W + dx (j + dx + 4 in 0 .. 11)
__device__ __forceinline__ float swin_at(const ScaleWin<true>& sw, int dy, int c) { return sw.g[dy][c >> 2][c & 3]; }
__global__ __launch_bounds__(kBlock) void k_svec(const float* __restrict__ A, size_t pitch, int N,
                                                 float* __restrict__ svec) {
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < N; i += (long)gridDim.x * kBlock) {
        const float d = A[(size_t)12 * pitch + i];
        svec[i] = d > 0.f ? 1.0f / sqrtf(d) : 1.0f;
    }
}

template <int FMT, bool SYM>
__global__ __launch_bounds__(kBlock) void k_scale16(const float* __restrict__ A, size_t pitch, int W, int N,
                                                    unsigned short* __restrict__ A16, size_t pitch16,
                                                    const float* __restrict__ svec) {
    // both products rounded to float, then the float to 16 bits: without
    // this the compiler fuses the second product and the conversion into one
    // mixed FMA (a single rounding), which differs where the float product
    // lands on a 16-bit rounding tie
#pragma clang fp contract(off)
    constexpr int K0 = SYM ? 13 : 0, NB = 25 - K0 - (SYM ? 0 : 1);
    const long nq = ((long)N + 3) / 4;
    const __amdgpu_buffer_rsrc_t bs = buf_rsrc(svec, N);
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < nq; q += (long)gridDim.x * kBlock) {
        const int ib = (int)(4 * q);
        float si[4], a[NB][4], sn[NB][4];
        ScaleWin<SYM> sw;
        if constexpr (SYM) load_swin(bs, W, ib, sw);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if constexpr (SYM) si[j] = swin_at(sw, 0, j + 4);
            else si[j] = ib + j < N ? svec[ib + j] : 0.f;
        }
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            const int k = K0 + m + (!SYM && K0 + m >= 12 ? 1 : 0);
            const int o = (k / 5 - 2) * W + (k % 5 - 2);
            // rows ib .. ib + 3 lie inside the padded pitch; 16-byte aligned
            const f32x4 v = *reinterpret_cast<const f32x4*>(A + (size_t)k * pitch + ib);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                a[m][j] = v[j];
                if constexpr (SYM) {
                    sn[m][j] = swin_at(sw, k / 5 - 2, j + k % 5 - 2 + 4);
                } else {
                    const long n = (long)ib + j + o;
                    sn[m][j] = (n >= 0 && n < N) ? svec[n] : 0.f;
                }
            }
        }
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            const int k = K0 + m + (!SYM && K0 + m >= 12 ? 1 : 0);
            unsigned int h[2];
#pragma unroll
            for (int j = 0; j < 4; j += 2) {
                const float v0 = (a[m][j] * si[j]) * sn[m][j];
                const float v1 = (a[m][j + 1] * si[j + 1]) * sn[m][j + 1];
                const unsigned int h0 = FMT == kF16 ? f32_to_f16_rne(v0) : f32_to_bf16_rne(v0);
                const unsigned int h1 = FMT == kF16 ? f32_to_f16_rne(v1) : f32_to_bf16_rne(v1);
                h[j / 2] = h0 | (h1 << 16);
            }
            *reinterpret_cast<u32x2*>(A16 + (size_t)k * pitch16 + ib) = u32x2{h[0], h[1]};
        }
    }
}

// The FP32 form of the symmetric scaled operator (the fused FP32 iteration):
// the 12 upper bands (a_ik s_i) s_{i+o} of As = S A S, band 13 + m at
// U + m * upitch, 0 where i + o_k leaves the vector.
// Each thread 4 consecutive rows (W % 4 == 0, N % 4 == 0), the neighbour
// scales as in k_scale16<*, true>.
__global__ __launch_bounds__(kBlock) void k_scale32(const float* __restrict__ A, size_t pitch, int W, int N,
                                                    float* __restrict__ U, size_t upitch,
                                                    const float* __restrict__ svec) {
#pragma clang fp contract(off)
    const __amdgpu_buffer_rsrc_t bs = buf_rsrc(svec, N);
    const long nq = (long)N / 4;
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < nq; q += (long)gridDim.x * kBlock) {
        const int ib = (int)(4 * q);
        ScaleWin<true> sw;
        load_swin(bs, W, ib, sw);
        f32x4 a[12];
#pragma unroll
        for (int m = 0; m < 12; ++m) a[m] = *reinterpret_cast<const f32x4*>(A + (size_t)(13 + m) * pitch + ib);
#pragma unroll
        for (int m = 0; m < 12; ++m) {
            const int k = 13 + m;
            f32x4 u;
#pragma unroll
            for (int j = 0; j < 4; ++j) u[j] = (a[m][j] * swin_at(sw, 0, j + 4)) * swin_at(sw, k / 5 - 2, j + k % 5 + 2);
            *reinterpret_cast<f32x4*>(U + (size_t)m * upitch + ib) = u;
        }
    }
}

// Gradient descent (main.cu:424-449): d = A x - b, <d,b>, ||d||^2
template <int R>
__global__ __launch_bounds__(kBlock) void k_gd1_25(const float* __restrict__ A, size_t pitch, int W, int N,
                                                   const float* __restrict__ x, const float* __restrict__ b,
                                                   float* __restrict__ d, Scalars* sc, double* partials,
                                                   unsigned int* counter) {
    if (sc->done != kRun) return;
    __shared__ __attribute__((aligned(16))) float lds[Tile<R>::LDS_FLOATS];
    const int ntiles = (N + Tile<R>::ROWS - 1) / Tile<R>::ROWS;
    const TileRange tr = xcd_tiles(ntiles);
    double v[2] = {0.0, 0.0};
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int i0 = t * Tile<R>::ROWS;
        __syncthreads();
        stage5<R>(lds, i0, W, N, [&](int g) { return x[g]; });
        __syncthreads();
        const int ib = i0 + R * threadIdx.x;
        if (ib < N) {
            float y[R], bb[R];
            tile_rows<R, kF32, kDiagStored>(A, pitch, ib, lds + R * threadIdx.x, y);
            load_rows<R>(b, ib, N, bb);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                y[j] -= bb[j];
                if (ib + j < N) {
                    v[0] += (double)(y[j] * bb[j]);
                    v[1] += (double)(y[j] * y[j]);
                }
            }
            store_rows<R>(d, ib, N, y);
        }
    }
    grid_reduce<2>(v, partials, counter, [&](double* tt) {
        sc->g0 = tt[0];
        sc->rr = tt[1];
        if (!isfinite(tt[0]) || !isfinite(tt[1])) sc->done = kNonFinite;
        else if (sc->tol2 >= 0 && tt[1] <= sc->tol2) sc->done = kConverged;
    });
}

// e = A d (not stored), <e,x>, <e,d> -> alpha = (<d,b> - <e,x>)/<e,d>
template <int R>
__global__ __launch_bounds__(kBlock) void k_gd2_25(const float* __restrict__ A, size_t pitch, int W, int N,
                                                   const float* __restrict__ x, const float* __restrict__ d,
                                                   Scalars* sc, double* partials, unsigned int* counter) {
    if (sc->done != kRun) return;
    __shared__ __attribute__((aligned(16))) float lds[Tile<R>::LDS_FLOATS];
    const int ntiles = (N + Tile<R>::ROWS - 1) / Tile<R>::ROWS;
    const TileRange tr = xcd_tiles(ntiles);
    double v[2] = {0.0, 0.0};
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int i0 = t * Tile<R>::ROWS;
        __syncthreads();
        stage5<R>(lds, i0, W, N, [&](int g) { return d[g]; });
        __syncthreads();
        const int ib = i0 + R * threadIdx.x;
        if (ib < N) {
            float y[R], xx[R];
            tile_rows<R, kF32, kDiagStored>(A, pitch, ib, lds + R * threadIdx.x, y);
            load_rows<R>(x, ib, N, xx);
#pragma unroll
            for (int j = 0; j < R; ++j)
                if (ib + j < N) {
                    v[0] += (double)(y[j] * xx[j]);
                    v[1] += (double)(y[j] * lds[2 * Tile<R>::SEG + R * threadIdx.x + 2 + j]);
                }
        }
    }
    grid_reduce<2>(v, partials, counter, [&](double* tt) {
        float k = (float)sc->g0;
        k = k - (float)tt[0];
        k = k / (float)tt[1];
        if (!isfinite(k)) {
            sc->done = tt[1] == 0 ? kConverged : kNonFinite;
            return;
        }
        sc->alpha = k;
    });
}

// x += alpha d; counts the iteration
__global__ __launch_bounds__(kBlock) void k_gd3(float* __restrict__ x, const float* __restrict__ d, int N, Scalars* sc,
                                                double* partials, unsigned int* counter) {
    if (sc->done != kRun) return;
    const float a = sc->alpha;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) x[i] = fmaf(a, d[i], x[i]);
    double v[1] = {0.0};
    grid_reduce<1>(v, partials, counter, [&](double*) {
        sc->iters += 1;
        if (sc->iters >= sc->maxit) sc->done = kMaxIt;
    });
}

// ---------------------------------------------------------------------------
// generic kernels (any band set, one row per thread, padded vectors)

template <int PM>
__global__ __launch_bounds__(kBlock) void k_resid_g(const float* __restrict__ a, size_t pitch, BandArgs ba, int N,
                                                    const float* __restrict__ x, const float* __restrict__ b,
                                                    const float* __restrict__ dinv, float* __restrict__ r, Scalars* sc,
                                                    double* partials, unsigned int* counter, int start) {
    if (!start && sc->done != kRun) return;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        float s = 0.f;
        for (int k = 0; k < ba.nb; ++k) s = fmaf(a[(size_t)i + (size_t)k * pitch], x[(long long)i + ba.off[k]], s);
        s -= b[i];
        r[i] = s;
        const float z = PM == kJacobi ? dinv[i] * s : s;
        v[0] += (double)(s * z);
        v[1] += (double)(s * s);
    }
    grid_reduce<2>(v, partials, counter, [&](double* tt) {
        sc->g0 = -1.0;  // an FP32 residual: no FP64 true residual of this x yet
        fin_residual(sc, tt[0], tt[1], true, false);
    });
}

template <int PM>
__global__ __launch_bounds__(kBlock) void k_cgstep_g(const float* __restrict__ a, size_t pitch, BandArgs ba, int N,
                                                     const float* __restrict__ r, const float* __restrict__ dinv,
                                                     float* __restrict__ p0, float* __restrict__ p1,
                                                     float* __restrict__ Ap, Scalars* sc, double* __restrict__ dpa,
                                                     const double* __restrict__ dpb, int verify) {
    StepIn in;
    if (!step_begin(sc, dpb, verify, in)) return;
    const float beta = in.beta;
    const int par = in.iters & 1;
    const float* __restrict__ pold = par ? p1 : p0;
    float* __restrict__ pnew = par ? p0 : p1;
    double v[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        float s = 0.f;
        for (int k = 0; k < ba.nb; ++k) {
            const long long g = (long long)i + ba.off[k];
            const float z = PM == kJacobi ? dinv[g] * r[g] : r[g];
            const float pg = beta != 0.f ? fmaf(beta, pold[g], -z) : -z;
            s = fmaf(a[(size_t)i + (size_t)k * pitch], pg, s);
        }
        const float zi = PM == kJacobi ? dinv[i] * r[i] : r[i];
        const float pi = beta != 0.f ? fmaf(beta, pold[i], -zi) : -zi;
        pnew[i] = pi;
        Ap[i] = s;
        v[0] += (double)(pi * s);
    }
    step_end(sc, v[0], dpa, in);
}

template <int PM, class XT>
__global__ __launch_bounds__(kBlock) void k_trueres_g(const float* __restrict__ a, size_t pitch, BandArgs ba, int N,
                                                      const XT* __restrict__ x, const float* __restrict__ b,
                                                      const float* __restrict__ dinv, float* __restrict__ r, int op,
                                                      Scalars* sc, double* partials, unsigned int* counter) {
    if (op != 0 && sc->done != kRun) return;
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        double s = 0.0;
        for (int k = 0; k < ba.nb; ++k)
            s = fma((double)a[(size_t)i + (size_t)k * pitch], (double)x[(long long)i + ba.off[k]], s);
        const double t = s - (double)b[i];
        const float rj = (float)t;
        if (op == 1) v[2] += (double)(rj - r[i]) * (double)(rj - r[i]);
        if (op != 0) r[i] = rj;
        v[1] += t * t;
        v[0] += PM == kJacobi ? (double)(rj * (dinv[i] * rj)) : (double)(rj * rj);
    }
    grid_reduce<3>(v, partials, counter, [&](double* tt) {
        sc->g0 = tt[1];
        if (op == 0) return;
        fin_residual(sc, tt[0], tt[1], op == 2, true, tt[2], tt[1]);
    });
}

__global__ __launch_bounds__(kBlock) void k_gd1_g(const float* __restrict__ a, size_t pitch, BandArgs ba, int N,
                                                  const float* __restrict__ x, const float* __restrict__ b,
                                                  float* __restrict__ d, Scalars* sc, double* partials,
                                                  unsigned int* counter) {
    if (sc->done != kRun) return;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        float s = 0.f;
        for (int k = 0; k < ba.nb; ++k) s = fmaf(a[(size_t)i + (size_t)k * pitch], x[(long long)i + ba.off[k]], s);
        s -= b[i];
        d[i] = s;
        v[0] += (double)(s * b[i]);
        v[1] += (double)(s * s);
    }
    grid_reduce<2>(v, partials, counter, [&](double* tt) {
        sc->g0 = tt[0];
        sc->rr = tt[1];
        if (!isfinite(tt[0]) || !isfinite(tt[1])) sc->done = kNonFinite;
        else if (sc->tol2 >= 0 && tt[1] <= sc->tol2) sc->done = kConverged;
    });
}

__global__ __launch_bounds__(kBlock) void k_gd2_g(const float* __restrict__ a, size_t pitch, BandArgs ba, int N,
                                                  const float* __restrict__ x, const float* __restrict__ d,
                                                  Scalars* sc, double* partials, unsigned int* counter) {
    if (sc->done != kRun) return;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        float s = 0.f;
        for (int k = 0; k < ba.nb; ++k) s = fmaf(a[(size_t)i + (size_t)k * pitch], d[(long long)i + ba.off[k]], s);
        v[0] += (double)(s * x[i]);
        v[1] += (double)(s * d[i]);
    }
    grid_reduce<2>(v, partials, counter, [&](double* tt) {
        float k = (float)sc->g0;
        k = k - (float)tt[0];
        k = k / (float)tt[1];
        if (!isfinite(k)) {
            sc->done = tt[1] == 0 ? kConverged : kNonFinite;
            return;
        }
        sc->alpha = k;
    });
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement (refine_mode 1)

// t = b - A x64 (FP64 accumulation, x64 padded), rhs = (float)t, g0 = ||t||^2
__global__ __launch_bounds__(kBlock) void k_refres64(const float* __restrict__ a, size_t pitch, BandArgs ba, int N,
                                                     const double* __restrict__ x64, const float* __restrict__ b,
                                                     float* __restrict__ rhs, Scalars* sc, double* partials,
                                                     unsigned int* counter) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        double s = 0.0;
        for (int k = 0; k < ba.nb; ++k) s = fma((double)a[(size_t)i + (size_t)k * pitch], x64[(long long)i + ba.off[k]], s);
        const double t = (double)b[i] - s;
        rhs[i] = (float)t;
        v[0] += t * t;
    }
    grid_reduce<1>(v, partials, counter, [&](double* tt) { sc->g0 = tt[0]; });
}

__global__ __launch_bounds__(kBlock) void k_zero64(double* v, int n) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[i] = 0.0;
}

// op 0: x64 = x (init); 1: x64 += x (accumulate a correction); 2: x = (float)x64
// (finish); 3: x64 += x, x = 0 (move the FP32 partial sum into the master);
// 4: x64 = x, x = 0 (start from the initial guess); 5: x64 += x, x = (float)x64
__global__ __launch_bounds__(kBlock) void k_x64(double* __restrict__ x64, float* __restrict__ x, int N, int op) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        if (op == 0) {
            x64[i] = (double)x[i];
        } else if (op == 1) {
            x64[i] += (double)x[i];
        } else if (op == 2) {
            x[i] = (float)x64[i];
        } else if (op == 5) {
            const double v = x64[i] + (double)x[i];
            x64[i] = v;
            x[i] = (float)v;
        } else {
            x64[i] = (op == 3 ? x64[i] : 0.0) + (double)x[i];
            x[i] = 0.f;
        }
    }
}

__global__ void k_tune_reset(Scalars* sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->done = 0;
        sc->verify = 0;
        sc->iters = 0;
        sc->pend = 0;
        sc->beta = 0.5f;
        sc->rz = 1.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill(float* v, int n, float val) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[i] = val;
}

__global__ __launch_bounds__(kBlock) void k_diag_inv(const float* __restrict__ diag, float* __restrict__ dinv, int N) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
        const float d = diag[i];
        dinv[i] = d != 0.f ? 1.f / d : 0.f;
    }
}

// ---------------------------------------------------------------------------
// workspaces: vectors padded by `pad` zeros (generic kernels read p, r and
// dinv at i + off), one allocation, each vector 256-byte aligned.

struct Workspace {
    int dev = -1;
    long n = 0;
    int pad = 0;
    float* base = nullptr;
    float *r = nullptr, *p0 = nullptr, *p1 = nullptr, *Ap = nullptr, *dinv = nullptr, *d = nullptr;
    unsigned short* A16 = nullptr;
    size_t pitch16 = 0;
    long cap16 = 0;
    double* x64 = nullptr;  // FP64 master x of the refinement mode (padded)
    long cap64 = 0;
    int pad64 = 0;
    Scalars* sc = nullptr;
    double* dpart = nullptr;  // deferred partial sums of the CG kernels: 3 x kMaxGrid, then 2 x 5 x kMaxGrid (fused)
    float *fbase = nullptr, *r2 = nullptr, *Ap2 = nullptr;  // second r and Ap of the fused iteration
    long capf = 0;
    unsigned int* uedge = nullptr;  // side words of the 16-bit operator's 256-wide columns (k_uedge)
    long cape = 0;
    Scalars* hsc = nullptr;  // pinned, 2 slots
    hipEvent_t ev[2] = {nullptr, nullptr};
};

static std::mutex g_ws_mu;
static std::map<std::pair<int, gm_stream_t>, Workspace*> g_ws;

static void ws_free_vectors(Workspace* w) {
    if (w->base) (void)hipFree(w->base);
    w->base = nullptr;
    w->n = 0;
}

static gm_status ws_get(gm_stream_t s, long n, int pad, bool need16, Workspace** out) {
    int dev = 0;
    GM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_ws_mu);
    Workspace*& w = g_ws[std::make_pair(dev, s)];
    if (!w) {
        w = new Workspace();
        w->dev = dev;
        GM_HIP(hipMalloc(&w->sc, sizeof(Scalars)));
        GM_HIP(hipMemset(w->sc, 0, sizeof(Scalars)));
        GM_HIP(hipMalloc(&w->dpart, sizeof(double) * 13 * kMaxGrid));
        GM_HIP(hipMemset(w->dpart, 0, sizeof(double) * 13 * kMaxGrid));
        GM_HIP(hipHostMalloc(&w->hsc, 2 * sizeof(Scalars), hipHostMallocDefault));
        GM_HIP(hipEventCreateWithFlags(&w->ev[0], hipEventDisableTiming));
        GM_HIP(hipEventCreateWithFlags(&w->ev[1], hipEventDisableTiming));
    }
    pad = (pad + 63) / 64 * 64;
    if (w->n < n || w->pad < pad) {
        ws_free_vectors(w);
        const long stride = (n + 2L * pad + 63) / 64 * 64;
        float* b = nullptr;
        GM_HIP(hipMalloc(&b, sizeof(float) * stride * 6));
        GM_HIP(hipMemset(b, 0, sizeof(float) * stride * 6));
        w->base = b;
        w->n = n;
        w->pad = pad;
        w->r = b + 0 * stride + pad;
        w->p0 = b + 1 * stride + pad;
        w->p1 = b + 2 * stride + pad;
        w->Ap = b + 3 * stride + pad;
        w->dinv = b + 4 * stride + pad;
        w->d = b + 5 * stride + pad;
    }
    if (need16 && w->cap16 < n) {
        if (w->A16) (void)hipFree(w->A16);
        w->A16 = nullptr;
        w->pitch16 = (size_t)((n + 63) / 64 * 64);
        GM_HIP(hipMalloc(&w->A16, sizeof(unsigned short) * w->pitch16 * 25 + 256));
        GM_HIP(hipMemset(w->A16, 0, sizeof(unsigned short) * w->pitch16 * 25 + 256));
        w->cap16 = n;
    }
    *out = w;
    return GM_OK;
}

// The fused iteration's second residual and product buffers (N floats each)
// and the edge array of a W x H operator.
static gm_status ws_fused(Workspace* w, long N, int W, int fmt) {
    const long ne = fused_edge_words(fmt, W, N);
    if (w->cape < ne) {
        if (w->uedge) (void)hipFree(w->uedge);
        w->uedge = nullptr;
        w->cape = 0;
        GM_HIP(hipMalloc(&w->uedge, sizeof(unsigned int) * ne));
        w->cape = ne;
    }
    if (w->capf < N) {
        if (w->fbase) (void)hipFree(w->fbase);
        w->fbase = w->r2 = w->Ap2 = nullptr;
        w->capf = 0;
        const long stride = (N + 63) / 64 * 64;
        GM_HIP(hipMalloc(&w->fbase, sizeof(float) * stride * 2));
        GM_HIP(hipMemset(w->fbase, 0, sizeof(float) * stride * 2));
        w->r2 = w->fbase;
        w->Ap2 = w->fbase + stride;
        w->capf = N;
    }
    return GM_OK;
}

// The FP64 master x of the refinement modes: N doubles padded by the largest
// band offset (the generic residual kernel reads x64[i + off]).  The pad past
// N is cleared here, since a larger earlier solve may have left values there.
static gm_status ws_x64(Workspace* w, long N, int maxoff, hipStream_t hs) {
    const int pad64 = (maxoff + 63) / 64 * 64;
    if (w->cap64 < N || w->pad64 < pad64) {
        if (w->x64) (void)hipFree(w->x64 - w->pad64);
        w->x64 = nullptr;
        w->cap64 = 0;
        double* p = nullptr;
        GM_HIP(hipMalloc(&p, sizeof(double) * ((size_t)N + 2 * pad64)));
        GM_HIP(hipMemset(p, 0, sizeof(double) * ((size_t)N + 2 * pad64)));
        w->x64 = p + pad64;
        w->cap64 = N;
        w->pad64 = pad64;
    }
    hipLaunchKernelGGL(k_zero64, dim3(1), dim3(kBlock), 0, hs, w->x64 + N, w->pad64);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

void release_workspaces() {
    std::lock_guard<std::mutex> g(g_ws_mu);
    for (auto& kv : g_ws) {
        Workspace* w = kv.second;
        ws_free_vectors(w);
        if (w->A16) (void)hipFree(w->A16);
        if (w->x64) (void)hipFree(w->x64 - w->pad64);
        if (w->fbase) (void)hipFree(w->fbase);
        if (w->uedge) (void)hipFree(w->uedge);
        if (w->sc) (void)hipFree(w->sc);
        if (w->dpart) (void)hipFree(w->dpart);
        if (w->hsc) (void)hipHostFree(w->hsc);
        for (int i = 0; i < 2; ++i)
            if (w->ev[i]) (void)hipEventDestroy(w->ev[i]);
        delete w;
    }
    g_ws.clear();
}

// ---------------------------------------------------------------------------
// host driver

// ---------------------------------------------------------------------------
// tile / grid choices of the fused CG step (K1) per format, filled by
// gm_autotune and consulted by every solve; defaults when untuned.
struct TuneKey {
    int fmt, W;
    long N;
    bool operator<(const TuneKey& o) const {
        return fmt != o.fmt ? fmt < o.fmt : (W != o.W ? W < o.W : N < o.N);
    }
};
struct TuneVal {
    int shape, per_cu;
    double us;
    int upd_pc;  // workgroups per CU of the update kernel (0 = default)
    int walk;    // symmetric 16-bit step: column-walk form (k_cgstep_symw)
};
static thread_local std::string g_tune_report;
static thread_local bool g_tune_spmv = false;  // the last gm_autotune on this thread also tuned the SpMV
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneVal> g_tune;

// the fused FP32 iteration's layout: the symmetric scaled operator of a
// W % 4 == 0 image, 4-row tiles of 128 columns
static bool fused32_ok(int W, long N) { return sym16_ok(W, N) && N < kFused32MaxN; }

static TuneVal tune_lookup(int fmt, int W, long N) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(TuneKey{fmt, W, N});
    if (it != g_tune.end()) return it->second;
    // defaults: the shapes the tuner picks for the 3840x2160 Levin matrix
    // FP32: the fused iteration on the symmetric FP32 form where its layout
    // applies (W % 4 == 0), else the Jacobi step + update
    if (fmt == kF32)
        return shape_ok(kS2R4Y4, fmt, W, N) ? TuneVal{kS2R4Y4, 2, 0, 0, fused32_ok(W, N) ? 2 : 0}
                                            : TuneVal{kS1R4, 4, 0, 0};
    return sym16_ok(W, N) ? TuneVal{kS2R4Y4, 2, 0, 0, 2} : TuneVal{kS1R8, 3, 0, 0, 0};
}

struct Solver {
    const BandedMatrix* A;
    BandPlan plan;
    bool tiled;
    int N;
    const float* b;
    float* x;
    Workspace* w;
    RedScratch rs;
    DevInfo dv;
    hipStream_t hs;
    int pm;        // PMode
    int fmt;       // kF32 / kF16 / kBF16
    int g_tile4, g_tile8, g_elem, g_gen;
    int shape, g_step;  // tuned K1 shape
    bool vec;           // x (and the workspace) 16-byte aligned: float4 update kernel
    int g_upd;
    int verify;         // recursive convergence asks for an FP64 check first
    bool sym;           // 16-bit form stored as its 12 upper bands (k_cgstep_sym)
    int walk;           // ... stepped by the column-walk kernel (k_cgstep_symw)
    double* x64;        // refine mode 2: FP64 master x (x then holds the FP32 partial sum since the last fold)
    bool fused;         // one launch per iteration (k_cgfused_symw)
    int fseq;           // fused launches so far (hand-off slot parity)
    int fstate;         // 0: the Scalars hold the state (next launch starts a direction); 1: the fused slots do
    int fdefer;         // the last fused launch deferred its update of x (mode 4)
};

static gm_status launch_resid(Solver& S, int start) {
    if (S.tiled) {
        if (S.pm == kJacobi)
            hipLaunchKernelGGL((k_resid25<4, kJacobi>), dim3(S.g_tile4), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.W, S.N, S.x, S.b, S.w->dinv, S.w->r, S.w->sc, S.rs.partials, S.rs.counter, start);
        else
            hipLaunchKernelGGL((k_resid25<4, kPlain>), dim3(S.g_tile4), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.W, S.N, S.x, S.b, S.w->dinv, S.w->r, S.w->sc, S.rs.partials, S.rs.counter, start);
    } else {
        if (S.pm == kJacobi)
            hipLaunchKernelGGL(k_resid_g<kJacobi>, dim3(S.g_gen), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.args, S.N, S.x, S.b, S.w->dinv, S.w->r, S.w->sc, S.rs.partials, S.rs.counter,
                               start);
        else
            hipLaunchKernelGGL(k_resid_g<kPlain>, dim3(S.g_gen), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.args, S.N, S.x, S.b, S.w->dinv, S.w->r, S.w->sc, S.rs.partials, S.rs.counter,
                               start);
    }
    GM_LAUNCH_CHECK();
    return GM_OK;
}

// xzero: the residual of the initial guess (k_bnorm recorded whether it is 0)
template <class XT>
static void launch_trueres_t(Solver& S, const XT* x, int op, int xzero) {
    using G = Geom2D<4, 4>;
    auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
    const bool seg = S.tiled && S.plan.W >= 4 && S.plan.W % 4 == 0 && S.N % S.plan.W == 0 &&
                     (double)S.A->apitch * 100.0 < 2147483648.0 && al16(x) && al16(S.b) && al16(S.w->dinv) &&
                     al16(S.w->r) && (!S.fused || al16(S.w->r2));
    if (seg) {
        // 768 x 4 tiles, 6 workgroups per CU (at 4K: 166.6 us, against 168.2 at 3 per CU and 183.3 us for
        // k_trueres2d's 256 x 4 tiles)
        constexpr int GR = 3;
        const int H = S.N / S.plan.W, ntiles = ((S.plan.W + 256 * GR - 1) / (256 * GR)) * ((H + 3) / 4);
        const int grid = grid_for_tiles(S.dv, ntiles, 6);
        float* r_odd = S.fused ? S.w->r2 : nullptr;
        if (S.pm == kJacobi)
            hipLaunchKernelGGL((k_trueres_seg<GR, kJacobi, XT>), dim3(grid), dim3(kBlock), 0, S.hs, S.A->a,
                               S.A->apitch, S.plan.W, H, S.N, x, S.b, S.w->dinv, S.w->r, r_odd, op, xzero, S.w->sc,
                               S.rs.partials);
        else if (S.pm == kScaled)
            hipLaunchKernelGGL((k_trueres_seg<GR, kScaled, XT>), dim3(grid), dim3(kBlock), 0, S.hs, S.A->a,
                               S.A->apitch, S.plan.W, H, S.N, x, S.b, S.w->dinv, S.w->r, r_odd, op, xzero, S.w->sc,
                               S.rs.partials);
        else
            hipLaunchKernelGGL((k_trueres_seg<GR, kPlain, XT>), dim3(grid), dim3(kBlock), 0, S.hs, S.A->a,
                               S.A->apitch, S.plan.W, H, S.N, x, S.b, S.w->dinv, S.w->r, r_odd, op, xzero, S.w->sc,
                               S.rs.partials);
        hipLaunchKernelGGL(k_trueres_fin, dim3(1), dim3(kBlock), 0, S.hs, S.w->sc, S.rs.partials, grid, op,
                           S.pm == kScaled ? 1 : 0);
    } else if (S.tiled && shape_ok(kS2R4Y4, kF32, S.plan.W, S.N)) {
        // 3 workgroups per CU (the FP32 step's pick), partial sums folded by k_trueres_fin
        const int grid = grid_for_tiles(S.dv, shape_tiles(kS2R4Y4, S.plan.W, S.N), 3);
        float* r_odd = S.fused ? S.w->r2 : nullptr;
        if (S.pm == kJacobi)
            hipLaunchKernelGGL((k_trueres2d<G, kJacobi, XT>), dim3(grid), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.W, S.N / S.plan.W, S.N, x, S.b, S.w->dinv, S.w->r, r_odd, op, xzero, S.w->sc,
                               S.rs.partials);
        else if (S.pm == kScaled)
            hipLaunchKernelGGL((k_trueres2d<G, kScaled, XT>), dim3(grid), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.W, S.N / S.plan.W, S.N, x, S.b, S.w->dinv, S.w->r, r_odd, op, xzero, S.w->sc,
                               S.rs.partials);
        else
            hipLaunchKernelGGL((k_trueres2d<G, kPlain, XT>), dim3(grid), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.W, S.N / S.plan.W, S.N, x, S.b, S.w->dinv, S.w->r, r_odd, op, xzero, S.w->sc,
                               S.rs.partials);
        hipLaunchKernelGGL(k_trueres_fin, dim3(1), dim3(kBlock), 0, S.hs, S.w->sc, S.rs.partials, grid, op,
                           S.pm == kScaled ? 1 : 0);
    } else if (S.tiled) {
        if (S.pm == kJacobi)
            hipLaunchKernelGGL((k_trueres25<4, kJacobi, XT>), dim3(S.g_tile4), dim3(kBlock), 0, S.hs, S.A->a,
                               S.A->apitch, S.plan.W, S.N, x, S.b, S.w->dinv, S.w->r, op, S.w->sc, S.rs.partials,
                               S.rs.counter);
        else if (S.pm == kScaled)
            hipLaunchKernelGGL((k_trueres25<4, kScaled, XT>), dim3(S.g_tile4), dim3(kBlock), 0, S.hs, S.A->a,
                               S.A->apitch, S.plan.W, S.N, x, S.b, S.w->dinv, S.w->r, op, S.w->sc, S.rs.partials,
                               S.rs.counter);
        else
            hipLaunchKernelGGL((k_trueres25<4, kPlain, XT>), dim3(S.g_tile4), dim3(kBlock), 0, S.hs, S.A->a,
                               S.A->apitch, S.plan.W, S.N, x, S.b, S.w->dinv, S.w->r, op, S.w->sc, S.rs.partials,
                               S.rs.counter);
    } else {
        if (S.pm == kJacobi)
            hipLaunchKernelGGL((k_trueres_g<kJacobi, XT>), dim3(S.g_gen), dim3(kBlock), 0, S.hs, S.A->a,
                               S.A->apitch, S.plan.args, S.N, x, S.b, S.w->dinv, S.w->r, op, S.w->sc, S.rs.partials,
                               S.rs.counter);
        else
            hipLaunchKernelGGL((k_trueres_g<kPlain, XT>), dim3(S.g_gen), dim3(kBlock), 0, S.hs, S.A->a,
                               S.A->apitch, S.plan.args, S.N, x, S.b, S.w->dinv, S.w->r, op, S.w->sc, S.rs.partials,
                               S.rs.counter);
    }
}

// The FP64-accumulated true residual of the current iterate (op as k_trueres25).
// Refine mode 2 first moves the FP32 partial sum into the FP64 master x, so the
// residual is that of the master and the FP32 x restarts from zero: x never
// carries more than the correction since the last fold, which removes the
// FP32 rounding floor of the true residual.
static gm_status fused_sync(Solver& S);

static gm_status launch_trueres(Solver& S, int op, int xzero = 0) {
    if (S.fused) {
        gm_status st = fused_sync(S);
        if (st != GM_OK) return st;
        S.fstate = 0;  // the next fused launch starts from this residual
    }
    if (S.x64) {
        hipLaunchKernelGGL(k_x64, dim3(S.g_elem), dim3(kBlock), 0, S.hs, S.x64, S.x, S.N, 3);
        launch_trueres_t<double>(S, S.x64, op, xzero);
    } else {
        launch_trueres_t<float>(S, S.x, op, xzero);
    }
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static void launch_scale16(Solver& S) {
    Workspace* w = S.w;
    const BandedMatrix* A = S.A;
    hipLaunchKernelGGL(k_svec, dim3(elementwise_grid(S.dv, S.N, 1)), dim3(kBlock), 0, S.hs, A->a, A->apitch, S.N,
                       w->dinv);
    const int grid = elementwise_grid(S.dv, ((long)S.N + 3) / 4, 1);
    if (S.fmt == kF32) {
        // the FP32 upper bands in the 16-bit operator's buffer (48 of its 50 bytes per row)
        hipLaunchKernelGGL(k_scale32, dim3(grid), dim3(kBlock), 0, S.hs, A->a, A->apitch,
                           S.plan.W, S.N, reinterpret_cast<float*>(w->A16), w->pitch16, w->dinv);
    } else if (S.fmt == kF16) {
        if (S.sym)
            hipLaunchKernelGGL((k_scale16<kF16, true>), dim3(grid), dim3(kBlock), 0, S.hs, A->a, A->apitch,
                               S.plan.W, S.N, w->A16, w->pitch16, w->dinv);
        else
            hipLaunchKernelGGL((k_scale16<kF16, false>), dim3(grid), dim3(kBlock), 0, S.hs, A->a, A->apitch,
                               S.plan.W, S.N, w->A16, w->pitch16, w->dinv);
    } else {
        if (S.sym)
            hipLaunchKernelGGL((k_scale16<kBF16, true>), dim3(grid), dim3(kBlock), 0, S.hs, A->a, A->apitch,
                               S.plan.W, S.N, w->A16, w->pitch16, w->dinv);
        else
            hipLaunchKernelGGL((k_scale16<kBF16, false>), dim3(grid), dim3(kBlock), 0, S.hs, A->a, A->apitch,
                               S.plan.W, S.N, w->A16, w->pitch16, w->dinv);
    }
}

// the fused iteration's edge array of the operator just scaled (columns of
// 256 for 16-bit bands, of 128 for FP32)
static void launch_uedge(Solver& S) {
    const int H = S.N / S.plan.W;
    const long total = fused_edge_words(S.fmt, S.plan.W, S.N);
    if (S.fmt == kF32)
        hipLaunchKernelGGL(k_uedge32, dim3(elementwise_grid(S.dv, total, 1)), dim3(kBlock), 0, S.hs,
                           reinterpret_cast<const float*>(S.w->A16), S.w->pitch16, S.plan.W, H, S.N, 128,
                           S.w->uedge);
    else
        hipLaunchKernelGGL(k_uedge, dim3(elementwise_grid(S.dv, total, 1)), dim3(kBlock), 0, S.hs, S.w->A16,
                           S.w->pitch16, S.plan.W, H, S.N, 256, S.w->uedge);
}

template <int TY, int FMT>
static void launch_sym_t(Solver& S) {
    Workspace* w = S.w;
    // the column walk's range-checked window loads address vectors of < 2^28 floats
    if (S.walk && S.N < (1 << 28))
        hipLaunchKernelGGL((k_cgstep_symw<TY, FMT>), dim3(S.g_step), dim3(kBlock), 0, S.hs, w->A16, w->pitch16,
                           S.plan.W, S.N / S.plan.W, S.N, w->r, w->p0, w->p1, w->Ap, w->sc, w->dpart,
                           w->dpart + kMaxGrid, S.verify);
    else
        hipLaunchKernelGGL((k_cgstep_sym<TY, FMT>), dim3(S.g_step), dim3(kBlock), 0, S.hs, w->A16, w->pitch16,
                           S.plan.W, S.N / S.plan.W, S.N, w->r, w->p0, w->p1, w->Ap, w->sc, w->dpart,
                           w->dpart + kMaxGrid, S.verify);
}

template <int FMT>
static void launch_sym(Solver& S) {
    switch (S.shape) {
        case kS2R4Y4: launch_sym_t<4, FMT>(S); break;
        case kS2R4Y16: launch_sym_t<16, FMT>(S); break;
        default: launch_sym_t<8, FMT>(S); break;
    }
}

template <class G, int FMT, int PM>
static void launch_step_t(Solver& S, const void* A, size_t pitch) {
    Workspace* w = S.w;
    hipLaunchKernelGGL((k_cgstep<G, FMT, PM>), dim3(S.g_step), dim3(kBlock), 0, S.hs, A, pitch, S.plan.W,
                       S.N / S.plan.W, S.N, w->r, w->dinv, w->p0, w->p1, w->Ap, w->sc, w->dpart,
                       w->dpart + kMaxGrid, S.verify);
}

// 16-bit forms without the symmetric layout (image width not a multiple of 4): 1-D tiles
template <int FMT>
static void launch_step_16(Solver& S, const void* A, size_t pitch) {
    if (S.shape == kS1R4) launch_step_t<Geom1D<4>, FMT, kScaled>(S, A, pitch);
    else launch_step_t<Geom1D<8>, FMT, kScaled>(S, A, pitch);
}

template <int FMT, int PM>
static void launch_step_shape(Solver& S, const void* A, size_t pitch) {
    switch (S.shape) {
        case kS1R2: launch_step_t<Geom1D<2>, FMT, PM>(S, A, pitch); break;
        case kS1R4: launch_step_t<Geom1D<4>, FMT, PM>(S, A, pitch); break;
        case kS1R8: launch_step_t<Geom1D<8>, FMT, PM>(S, A, pitch); break;
        case kS2R4Y4: launch_step_t<Geom2D<4, 4>, FMT, PM>(S, A, pitch); break;
        case kS2R4Y8: launch_step_t<Geom2D<4, 8>, FMT, PM>(S, A, pitch); break;
        case kS2R4Y16: launch_step_t<Geom2D<4, 16>, FMT, PM>(S, A, pitch); break;
        case kS2R8Y8: launch_step_t<Geom2D<8, 8>, FMT, PM>(S, A, pitch); break;
        default: launch_step_t<Geom2D<8, 16>, FMT, PM>(S, A, pitch); break;
    }
}

static gm_status launch_step(Solver& S) {
    Workspace* w = S.w;
    if (S.tiled) {
        if (S.pm == kScaled && S.sym) {
            if (S.fmt == kF16) launch_sym<kF16>(S);
            else launch_sym<kBF16>(S);
        } else if (S.pm == kScaled) {
            if (S.fmt == kF16) launch_step_16<kF16>(S, w->A16, w->pitch16);
            else launch_step_16<kBF16>(S, w->A16, w->pitch16);
        } else if (S.pm == kJacobi) {
            launch_step_shape<kF32, kJacobi>(S, S.A->a, S.A->apitch);
        } else {
            launch_step_shape<kF32, kPlain>(S, S.A->a, S.A->apitch);
        }
    } else {
        if (S.pm == kJacobi)
            hipLaunchKernelGGL(k_cgstep_g<kJacobi>, dim3(S.g_gen), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.args, S.N, w->r, w->dinv, w->p0, w->p1, w->Ap, w->sc, w->dpart,
                               w->dpart + kMaxGrid, S.verify);
        else
            hipLaunchKernelGGL(k_cgstep_g<kPlain>, dim3(S.g_gen), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch,
                               S.plan.args, S.N, w->r, w->dinv, w->p0, w->p1, w->Ap, w->sc, w->dpart,
                               w->dpart + kMaxGrid, S.verify);
    }
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static void set_step_shape(Solver& S, int shape, int per_cu) {
    S.shape = shape;
    S.g_step = grid_for_tiles(S.dv, shape_tiles(shape, S.plan.W, S.N), per_cu);
}

template <int PM, bool VEC>
static void launch_update_t(Solver& S, int grid) {
    Workspace* w = S.w;
    hipLaunchKernelGGL((k_cgupdate<PM, VEC>), dim3(grid), dim3(kBlock), 0, S.hs, S.x, w->r, w->dinv, w->p0, w->p1,
                       w->Ap, S.N, w->sc, w->dpart, w->dpart + kMaxGrid);
}

static gm_status launch_update(Solver& S) {
    if (S.vec) {
        if (S.pm == kJacobi) launch_update_t<kJacobi, true>(S, S.g_upd);
        else if (S.pm == kScaled) launch_update_t<kScaled, true>(S, S.g_upd);
        else launch_update_t<kPlain, true>(S, S.g_upd);
    } else {
        if (S.pm == kJacobi) launch_update_t<kJacobi, false>(S, S.g_elem);
        else if (S.pm == kScaled) launch_update_t<kScaled, false>(S, S.g_elem);
        else launch_update_t<kPlain, false>(S, S.g_elem);
    }
    GM_LAUNCH_CHECK();
    return GM_OK;
}

template <int FMT>
static void launch_fused_t(Solver& S, int mode) {
    Workspace* w = S.w;
    const void* up = FMT == kF32 ? static_cast<const void*>(w->A16) : static_cast<const void*>(w->A16 + 13 * w->pitch16);
    // the form: deferring (modes 4, 5), making a deferred update too (mode 2,
    // and modes 0 and 3 after a deferring launch), or neither
    const int xm = (mode == 4 || mode == 5) ? kXDefer
                   : (mode == 2 || ((mode == 0 || mode == 3) && S.fdefer)) ? kXFlush
                                                                          : kXPlain;
    auto k = xm == kXDefer ? k_cgfused_symw<FMT, kXDefer>
             : xm == kXFlush ? k_cgfused_symw<FMT, kXFlush>
                             : k_cgfused_symw<FMT, kXPlain>;
    hipLaunchKernelGGL(k, dim3(S.g_step), dim3(kBlock), 0, S.hs, up, w->pitch16, S.plan.W, S.N / S.plan.W, S.N,
                       w->dinv, S.x, w->r, w->r2, w->p0, w->p1, w->Ap, w->Ap2, w->uedge, w->sc,
                       w->dpart + 3 * kMaxGrid, S.fseq, mode, S.verify);
    ++S.fseq;
}

// one fused launch (mode as k_cgfused_symw)
static gm_status launch_fused(Solver& S, int mode) {
    if (S.fmt == kF16) launch_fused_t<kF16>(S, mode);
    else if (S.fmt == kBF16) launch_fused_t<kBF16>(S, mode);
    else launch_fused_t<kF32>(S, mode);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

// Publish the fused iteration's state into the Scalars (nothing to do while
// they hold it already).
static gm_status fused_sync(Solver& S) {
    if (!S.fused || !S.fstate) return GM_OK;
    hipLaunchKernelGGL(k_fused_fold, dim3(1), dim3(kBlock), 0, S.hs, S.w->sc, S.w->dpart + 3 * kMaxGrid, S.fseq,
                       S.verify);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

// replace_next: a residual from scratch follows this iteration.  may_defer:
// the next launch is another iteration (no fold, no replacement, below the
// iteration cap), which may then make this one's update of x.
static gm_status launch_cg_iter(Solver& S, bool replace_next = false, bool may_defer = false) {
    gm_status st;
    if (S.fused) {
        // a residual from scratch first takes its step (the first direction
        // or the one kept across a replacement), then one launch per
        // iteration; the iteration before a replacement makes only its
        // update; every other launch defers its update of x to the next
        if (!S.fstate) {
            if ((st = launch_fused(S, 1)) != GM_OK) return st;
            S.fstate = 1;
            S.fdefer = 0;
        }
        const int mode = replace_next ? 3 : (may_defer && !S.fdefer) ? 4 : 0;
        st = launch_fused(S, mode);  // (the form after a deferring launch: S.fdefer)
        S.fdefer = mode == 4;
        return st;
    }
    if ((st = launch_step(S)) != GM_OK) return st;
    return launch_update(S);
}

// fold the last update's pending partial sums into the Scalars
static gm_status launch_fold(Solver& S) {
    if (S.fused) return fused_sync(S);
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(kBlock), 0, S.hs, S.w->sc, S.w->dpart + kMaxGrid, S.verify);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status launch_gd_iter(Solver& S) {
    Workspace* w = S.w;
    if (S.tiled) {
        hipLaunchKernelGGL((k_gd1_25<4>), dim3(S.g_tile4), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch, S.plan.W, S.N,
                           S.x, S.b, w->d, w->sc, S.rs.partials, S.rs.counter);
        GM_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_gd2_25<4>), dim3(S.g_tile4), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch, S.plan.W, S.N,
                           S.x, w->d, w->sc, S.rs.partials, S.rs.counter);
    } else {
        hipLaunchKernelGGL(k_gd1_g, dim3(S.g_gen), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch, S.plan.args, S.N, S.x,
                           S.b, w->d, w->sc, S.rs.partials, S.rs.counter);
        GM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gd2_g, dim3(S.g_gen), dim3(kBlock), 0, S.hs, S.A->a, S.A->apitch, S.plan.args, S.N, S.x,
                           w->d, w->sc, S.rs.partials, S.rs.counter);
    }
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gd3, dim3(S.g_elem), dim3(kBlock), 0, S.hs, S.x, w->d, S.N, w->sc, S.rs.partials,
                       S.rs.counter);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status status_of_done(int done) {
    switch (done) {
        case kConverged: return GM_OK;
        case kBreakdown: return GM_ERR_BREAKDOWN;
        case kNonFinite: return GM_ERR_NONFINITE;
        case kMaxIt: return GM_ERR_NOT_CONVERGED;
        default: return GM_OK;
    }
}

gm_status solve_impl(float* x, const BandedMatrix* A, const float* b, int pad, const gm_solve_opts* o,
                     gm_solve_info* info, gm_stream_t s);

// Iterative refinement with an FP64 master x: r = b - A x64 (FP64), inner
// solve A d = r with the requested method/precision to a relative 1e-4
// (recursive residual), x64 += d; until ||b - A x64|| <= tol ||b||.
static gm_status refine_impl(float* x, const BandedMatrix* A, const float* b, int pad, const gm_solve_opts* o,
                             gm_solve_info* info, gm_stream_t s) {
    BandPlan plan;
    gm_status st = get_plan(A, &plan);
    if (st != GM_OK) return st;
    if (!x || !b) return fail(GM_ERR_INVALID, "solve: null vector");
    const int N = A->rows;
    if (info) {
        memset(info, 0, sizeof(*info));
        info->true_relres = -1;
    }
    if (N == 0) return GM_OK;
    DevInfo dv;
    RedScratch rs;
    Workspace* w;
    if ((st = dev_info(&dv)) != GM_OK) return st;
    if ((st = red_scratch(s, &rs)) != GM_OK) return st;
    if ((st = ws_get(s, N, plan.maxoff, o->precision != GM_PREC_FP32, &w)) != GM_OK) return st;
    hipStream_t hs = (hipStream_t)s;
    if ((st = ws_x64(w, N, plan.maxoff, hs)) != GM_OK) return st;
    const int ge = elementwise_grid(dv, N, 4);
    hipLaunchKernelGGL(k_x64, dim3(ge), dim3(kBlock), 0, hs, w->x64, x, N, 0);
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, hs, w->sc, 1);
    launch_bnorm(dv, hs, b, nullptr, N, 0.0, w->sc, rs.partials);
    GM_LAUNCH_CHECK();
    Scalars h;
    GM_HIP(hipMemcpyAsync(&h, w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, hs));
    GM_HIP(hipStreamSynchronize(hs));
    const double bb = h.bb > 0 ? h.bb : 1.0;
    gm_solve_opts in = *o;
    in.refine_mode = 0;
    in.verify = 0;
    in.tol = 1e-4;
    int iters = 0, outer = 0;
    double relres = 0;
    gm_status result = GM_ERR_NOT_CONVERGED;
    float* rhs = w->d;
    double prev = 1e300;
    for (outer = 0; outer < 30; ++outer) {
        hipLaunchKernelGGL(k_refres64, dim3(elementwise_grid(dv, N, 1)), dim3(kBlock), 0, hs, A->a, A->apitch,
                           plan.args, N, w->x64, b, rhs, w->sc, rs.partials, rs.counter);
        GM_LAUNCH_CHECK();
        GM_HIP(hipMemcpyAsync(&h, w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, hs));
        GM_HIP(hipStreamSynchronize(hs));
        relres = sqrt(h.g0 / bb);
        if (!std::isfinite(relres)) {
            result = GM_ERR_NONFINITE;
            break;
        }
        if (relres <= o->tol) {
            result = GM_OK;
            break;
        }
        if (iters >= o->max_iter || relres > 0.9 * prev) break;  // budget spent / no progress
        prev = relres;
        in.max_iter = o->max_iter - iters;
        hipLaunchKernelGGL(k_fill, dim3(ge), dim3(kBlock), 0, hs, x, N, 0.0f);
        gm_solve_info ii;
        st = solve_impl(x, A, rhs, pad, &in, &ii, s);
        iters += ii.iterations;
        if (st != GM_OK && st != GM_ERR_NOT_CONVERGED) {
            result = st;
            break;
        }
        hipLaunchKernelGGL(k_x64, dim3(ge), dim3(kBlock), 0, hs, w->x64, x, N, 1);
        GM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_x64, dim3(ge), dim3(kBlock), 0, hs, w->x64, x, N, 2);
    GM_LAUNCH_CHECK();
    GM_HIP(hipStreamSynchronize(hs));
    if (info) {
        info->iterations = iters;
        info->status = result;
        info->relres = relres;
        info->true_relres = relres;
        info->replacements = outer;
    }
    if (result == GM_ERR_NOT_CONVERGED)
        set_error("refinement: relres %.3e after %d outer steps / %d iterations", relres, outer, iters);
    return result;
}

gm_status solve_impl(float* x, const BandedMatrix* A, const float* b, int pad, const gm_solve_opts* o,
                     gm_solve_info* info, gm_stream_t s) {
    if (o->refine_mode == 1) return refine_impl(x, A, b, pad, o, info, s);
    if (o->refine_mode < 0 || o->refine_mode > 2) return fail(GM_ERR_INVALID, "solve: refine_mode");
    Solver S{};
    gm_status st = get_plan(A, &S.plan);
    if (st != GM_OK) return st;
    if (!x || !b) return fail(GM_ERR_INVALID, "solve: null vector");
    if (o->max_iter <= 0) return fail(GM_ERR_INVALID, "solve: max_iter must be > 0");
    if (o->method < GM_METHOD_CG || o->method > GM_METHOD_GD) return fail(GM_ERR_INVALID, "solve: method");
    if (o->precision < GM_PREC_FP32 || o->precision > GM_PREC_BF16) return fail(GM_ERR_INVALID, "solve: precision");
    S.A = A;
    S.N = A->rows;
    S.b = b;
    S.x = x;
    S.hs = (hipStream_t)s;
    S.tiled = tiled_ok(S.plan, A->a, A->apitch);
    if (!S.tiled && pad < S.plan.maxoff)
        return fail(GM_ERR_INVALID, "solve: pad %d < largest band offset %d", pad, S.plan.maxoff);
    const bool mixed = o->precision != GM_PREC_FP32;
    if (mixed && (!S.tiled || o->method == GM_METHOD_GD))
        return fail(GM_ERR_UNSUPPORTED, "16-bit forms need the 5x5 Levin band structure and CG/PCG");
    if (o->refine_mode == 2 && o->method == GM_METHOD_GD)
        return fail(GM_ERR_UNSUPPORTED, "refine_mode 2 needs CG/PCG");
    if ((o->method == GM_METHOD_PCG || mixed) && S.plan.args.center < 0)
        return fail(GM_ERR_INVALID, "solve: Jacobi scaling needs a band with offset 0");
    S.fmt = o->precision == GM_PREC_FP16 ? kF16 : (o->precision == GM_PREC_BF16 ? kBF16 : kF32);
    const TuneVal tv = tune_lookup(S.fmt, S.plan.W, S.N);
    S.vec = (reinterpret_cast<uintptr_t>(x) & 15u) == 0;
    // FP32 Jacobi-PCG on a W % 4 == 0 Levin matrix, when the tuned (or
    // default) form is the fused one: CG on the symmetric scaled FP32 operator
    // S A S (upper bands), as the 16-bit forms, with the FP64 residual
    // replacement through A itself
    const bool f32fused = !mixed && o->method == GM_METHOD_PCG && S.tiled && S.N > 0 &&
                          fused32_ok(S.plan.W, S.N) && tv.walk == 2 && tv.shape == kS2R4Y4 && S.vec;
    S.pm = (mixed || f32fused) ? kScaled : (o->method == GM_METHOD_PCG ? kJacobi : kPlain);
    S.sym = (mixed && S.tiled && sym16_ok(S.plan.W, S.N)) || f32fused;
    if ((st = dev_info(&S.dv)) != GM_OK) return st;
    if ((st = red_scratch(s, &S.rs)) != GM_OK) return st;
    if ((st = ws_get(s, S.N > 0 ? S.N : 1, S.plan.maxoff, mixed || f32fused, &S.w)) != GM_OK) return st;
    if (info) {
        memset(info, 0, sizeof(*info));
        info->true_relres = -1;
    }
    if (S.N == 0) return GM_OK;
    S.g_tile4 = grid_for_tiles(S.dv, (S.N + Tile<4>::ROWS - 1) / Tile<4>::ROWS, 4);
    S.g_tile8 = grid_for_tiles(S.dv, (S.N + Tile<8>::ROWS - 1) / Tile<8>::ROWS, 3);
    S.g_elem = elementwise_grid(S.dv, S.N, 4);
    S.g_gen = elementwise_grid(S.dv, S.N, 1);
    S.g_upd = elementwise_grid(S.dv, (S.N + 3) / 4, 2);
    {
        if (S.tiled) set_step_shape(S, tv.shape, tv.per_cu);
        S.walk = tv.walk;
        // the fused iteration: symmetric 16-bit (or FP32) operator, 4-row
        // column walk, 16-byte aligned x
        S.fused = S.sym && tv.walk == 2 && S.shape == kS2R4Y4 && S.N < fused_max_n(S.fmt) && S.vec &&
                  o->method != GM_METHOD_GD;
        if (tv.upd_pc > 0) S.g_upd = grid_for_tiles(S.dv, (S.N + 4L * kBlock - 1) / (4L * kBlock), tv.upd_pc);
    }
    // the scaled FP32 operator exists only in the fused iteration's form
    if (S.fmt == kF32 && S.pm == kScaled && !S.fused) return fail(GM_ERR_INVALID, "solve: internal: FP32 scaled form");
    Workspace* w = S.w;
    if (S.fused && (st = ws_fused(w, S.N, S.plan.W, S.fmt)) != GM_OK) return st;
    if (!S.tiled) {
        // generic kernels read r, p and D^-1 at i + offset up to N + pad: clear
        // whatever a larger earlier solve left beyond N in this workspace
        for (float* v : {w->r, w->p0, w->p1, w->Ap, w->dinv, w->d})
            hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, S.hs, v + S.N, w->pad, 0.0f);
        GM_LAUNCH_CHECK();
    }

    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, S.hs, w->sc, o->max_iter);
    GM_LAUNCH_CHECK();
    launch_bnorm(S.dv, S.hs, b, x, S.N, o->tol, w->sc, S.rs.partials);
    GM_LAUNCH_CHECK();
    if (o->refine_mode == 2) {
        // the initial guess moves into the FP64 master; x accumulates from zero
        if ((st = ws_x64(w, S.N, S.plan.maxoff, S.hs)) != GM_OK) return st;
        S.x64 = w->x64;
        hipLaunchKernelGGL(k_x64, dim3(S.g_elem), dim3(kBlock), 0, S.hs, S.x64, x, S.N, 4);
        GM_LAUNCH_CHECK();
    }

    S.verify = (o->tol > 0 && o->verify) ? 1 : 0;
    if (S.pm == kJacobi) {
        hipLaunchKernelGGL(k_diag_inv, dim3(S.g_elem), dim3(kBlock), 0, S.hs, A->a + (size_t)S.plan.args.center * A->apitch,
                           w->dinv, S.N);
        GM_LAUNCH_CHECK();
    } else if (S.pm == kScaled) {
        launch_scale16(S);
        GM_LAUNCH_CHECK();
        if (S.fused) launch_uedge(S);
    }

    const int chunk = o->check_every > 0 ? o->check_every : 16;
    // FP64 residual replacement period: the scaled forms (16-bit, fused FP32)
    // and refine mode 2 default to 50, the Jacobi FP32 step to none
    const int refine =
        o->refine_every > 0 ? o->refine_every : (o->refine_every == 0 && (S.pm == kScaled || S.x64) ? 50 : 0);
    long enq = 0;  // iterations enqueued
    int polls = 0;
    bool finished = false;

    if (o->method == GM_METHOD_GD) {
        while (!finished) {
            for (int j = 0; j < chunk && enq < o->max_iter; ++j, ++enq)
                if ((st = launch_gd_iter(S)) != GM_OK) return st;
            GM_HIP(hipMemcpyAsync(&w->hsc[polls & 1], w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S.hs));
            GM_HIP(hipEventRecord(w->ev[polls & 1], S.hs));
            if (polls > 0) {
                GM_HIP(hipEventSynchronize(w->ev[(polls - 1) & 1]));
                if (w->hsc[(polls - 1) & 1].done != kRun) finished = true;
            }
            ++polls;
            if (enq >= o->max_iter) finished = true;
        }
    } else {
        if (S.pm == kScaled || S.x64) {
            if ((st = launch_trueres(S, 2, 1)) != GM_OK) return st;
        } else {
            if ((st = launch_resid(S, 1)) != GM_OK) return st;
        }
        long since_restart = 0;
        // Chunk size: `chunk` until the residual's decay rate (from the last
        // two polls) predicts convergence within the iterations already
        // queued; then small chunks, so that few launches find the solve
        // already stopped.
        int next = chunk;
        double rr_a = -1.0, rr_b = -1.0;
        long it_a = 0, it_b = 0;
        while (!finished) {
            const int this_chunk = next;
            for (int j = 0; j < this_chunk && enq < o->max_iter; ++j) {
                const bool restart_next = o->restart > 0 && since_restart + 1 >= o->restart && enq + 1 < o->max_iter;
                const bool refine_next = refine > 0 && (enq + 1) % refine == 0;
                const bool may_defer = j + 1 < this_chunk && enq + 1 < o->max_iter;
                if ((st = launch_cg_iter(S, restart_next || refine_next, may_defer)) != GM_OK) return st;
                ++enq;
                ++since_restart;
                bool restarted = false;
                if (o->restart > 0 && since_restart >= o->restart && enq < o->max_iter) {
                    // cgSolve's periodic restart (main.cu:483-493): r from scratch, p = -z
                    // (the scaled 16-bit system's residual is s (A x - b): the
                    // true-residual kernel forms it, as at the start)
                    if ((st = (S.pm == kScaled || S.x64) ? launch_trueres(S, 2) : launch_resid(S, 0)) != GM_OK)
                        return st;
                    since_restart = 0;
                    restarted = true;
                }
                // the FP64 residual replacement, unless a restart has just formed
                // the residual from scratch (a replacement keeps the direction,
                // and would undo the restart's p = -z)
                if (refine > 0 && enq % refine == 0 && !restarted) {
                    if ((st = launch_trueres(S, 1)) != GM_OK) return st;
                }
            }
            if ((st = fused_sync(S)) != GM_OK) return st;
            GM_HIP(hipMemcpyAsync(&w->hsc[polls & 1], w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S.hs));
            GM_HIP(hipEventRecord(w->ev[polls & 1], S.hs));
            if (polls > 0) {
                GM_HIP(hipEventSynchronize(w->ev[(polls - 1) & 1]));
                const Scalars& h = w->hsc[(polls - 1) & 1];
                if (h.done != kRun) {
                    finished = true;
                } else if (h.verify) {
                    // recursive residual reached the tolerance: check the true
                    // one, and wait for the verdict before queueing more
                    if ((st = launch_trueres(S, 1)) != GM_OK) return st;
                    GM_HIP(hipMemcpyAsync(&w->hsc[polls & 1], w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S.hs));
                    GM_HIP(hipStreamSynchronize(S.hs));
                    if (w->hsc[polls & 1].done != kRun) finished = true;
                    rr_a = rr_b = -1.0;
                    next = chunk;
                } else if (h.tol2 > 0 && h.rr > 0) {
                    rr_a = rr_b;
                    it_a = it_b;
                    rr_b = h.rr;
                    it_b = h.iters;
                    next = chunk;
                    if (rr_a > 0 && rr_b < rr_a && it_b > it_a) {
                        const double rate = log(rr_b / rr_a) / (double)(it_b - it_a);  // < 0 per iteration
                        const double left = log(h.tol2 / rr_b) / rate - (double)(enq - it_b);
                        if (left < chunk) next = left < 1.0 ? 2 : (int)ceil(left) + 1;
                        if (next > chunk) next = chunk;
                    }
                }
            }
            ++polls;
            if (enq >= o->max_iter) {
                // drain: fold the last update, then a pending verification may
                // still need the final check
                if ((st = launch_fold(S)) != GM_OK) return st;
                GM_HIP(hipMemcpyAsync(&w->hsc[polls & 1], w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S.hs));
                GM_HIP(hipStreamSynchronize(S.hs));
                const Scalars& h = w->hsc[polls & 1];
                if (h.done == kRun && h.verify) {
                    if ((st = launch_trueres(S, 1)) != GM_OK) return st;
                }
                finished = true;
            }
        }
    }
    if (o->method != GM_METHOD_GD && (st = launch_fold(S)) != GM_OK) return st;
    // the final state through the pinned slot, ordered on the solve's stream
    // (a synchronous copy into pageable memory idled the GPU ~0.2 ms per solve)
    GM_HIP(hipMemcpyAsync(&w->hsc[0], w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S.hs));
    GM_HIP(hipStreamSynchronize(S.hs));
    Scalars fin = w->hsc[0];
    const double bb = fin.bb > 0 ? fin.bb : 1.0;
    gm_status result = status_of_done(fin.done);
    if (fin.done == kRun) result = (o->tol > 0) ? GM_ERR_NOT_CONVERGED : GM_OK;
    if (o->tol <= 0 && fin.done == kMaxIt) result = GM_OK;  // fixed-iteration run
    if (info) {
        info->iterations = fin.iters;
        info->status = result;
        info->relres = sqrt(fin.rr / bb);
        info->replacements = fin.replacements;
        info->true_relres = -1;
    }
    if (o->verify && info && o->method != GM_METHOD_GD) {
        // FP64-accumulated true residual of the returned x.  With verification
        // on, convergence is only ever declared by a true-residual pass, after
        // which x is not updated again: its g0 is already that of the final x.
        if (!(S.verify && fin.done == kConverged && fin.g0 >= 0)) {
            if ((st = launch_trueres(S, 0)) != GM_OK) return st;
            GM_HIP(hipMemcpyAsync(&w->hsc[0], w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S.hs));
            GM_HIP(hipStreamSynchronize(S.hs));
            fin = w->hsc[0];
        }
        info->true_relres = sqrt(fin.g0 / bb);
    }
    if (S.x64) {
        // returned x = the FP64 master (with any partial sum not yet folded), rounded
        hipLaunchKernelGGL(k_x64, dim3(S.g_elem), dim3(kBlock), 0, S.hs, S.x64, x, S.N, 5);
        GM_LAUNCH_CHECK();
        GM_HIP(hipStreamSynchronize(S.hs));
    }
    if (result != GM_OK && result != GM_ERR_NOT_CONVERGED) {
        set_error("solve stopped: %s after %d iterations",
                  result == GM_ERR_BREAKDOWN ? "breakdown (p'Ap <= 0)" : "non-finite value", fin.iters);
    } else if (result == GM_ERR_NOT_CONVERGED) {
        set_error("solve: tolerance not reached in %d iterations (relres %.3e)", fin.iters, sqrt(fin.rr / bb));
    }
    return result;
}

}  // namespace gm

using namespace gm;

extern "C" {

void gm_solve_opts_default(gm_solve_opts* o) {
    if (!o) return;
    o->method = GM_METHOD_PCG;
    o->precision = GM_PREC_FP32;
    o->max_iter = 10000;
    o->tol = 1e-6;
    o->restart = 0;
    o->refine_every = 0;
    o->check_every = 16;
    o->verify = 1;
    o->refine_mode = 0;
}

gm_status gm_solve(float* alpha, const BandedMatrix* L, const float* b, int pad, const gm_solve_opts* opts,
                   gm_solve_info* info, gm_stream_t stream) {
    gm_solve_opts o;
    if (opts) o = *opts;
    else gm_solve_opts_default(&o);
    return solve_impl(alpha, L, b, pad, &o, info, stream);
}

gm_status gm_cg_solve(float* alpha, const BandedMatrix* L, const float* b, int pad, int iterations,
                      int restart_interval, gm_stream_t stream) {
    if (iterations <= 0) return GM_OK;
    gm_solve_opts o;
    gm_solve_opts_default(&o);
    o.method = GM_METHOD_CG;
    o.max_iter = iterations;
    o.tol = 0;  // fixed iteration count (main.cu:481)
    o.restart = restart_interval > 0 ? restart_interval : 0;
    o.verify = 0;
    o.check_every = 32;
    gm_status st = solve_impl(alpha, L, b, pad, &o, nullptr, stream);
    return st == GM_ERR_NOT_CONVERGED ? GM_OK : st;
}

gm_status gm_grad_solve(float* alpha, const BandedMatrix* L, const float* b, int iterations, int pad,
                        gm_stream_t stream) {
    if (iterations <= 0) return GM_OK;
    gm_solve_opts o;
    gm_solve_opts_default(&o);
    o.method = GM_METHOD_GD;
    o.max_iter = iterations;
    o.tol = 0;
    o.verify = 0;
    o.check_every = 32;
    gm_status st = solve_impl(alpha, L, b, pad, &o, nullptr, stream);
    return st == GM_ERR_NOT_CONVERGED ? GM_OK : st;
}

gm_status gm_true_relres(const BandedMatrix* A, const float* x, const float* b, double* relres, gm_stream_t s) {
    if (!relres) return fail(GM_ERR_INVALID, "null relres");
    Solver S{};
    gm_status st = get_plan(A, &S.plan);
    if (st != GM_OK) return st;
    if (!x || !b) return fail(GM_ERR_INVALID, "null vector");
    S.A = A;
    S.N = A->rows;
    S.b = b;
    S.x = const_cast<float*>(x);
    S.hs = (hipStream_t)s;
    S.tiled = tiled_ok(S.plan, A->a, A->apitch);
    S.pm = kPlain;
    if ((st = dev_info(&S.dv)) != GM_OK) return st;
    if ((st = red_scratch(s, &S.rs)) != GM_OK) return st;
    if ((st = ws_get(s, S.N > 0 ? S.N : 1, S.plan.maxoff, false, &S.w)) != GM_OK) return st;
    if (S.N == 0) {
        *relres = 0;
        return GM_OK;
    }
    S.g_tile4 = grid_for_tiles(S.dv, (S.N + Tile<4>::ROWS - 1) / Tile<4>::ROWS, 4);
    S.g_elem = elementwise_grid(S.dv, S.N, 4);
    S.g_gen = elementwise_grid(S.dv, S.N, 1);
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(64), 0, S.hs, S.w->sc, 1);
    GM_LAUNCH_CHECK();
    launch_bnorm(S.dv, S.hs, b, nullptr, S.N, 0.0, S.w->sc, S.rs.partials);
    GM_LAUNCH_CHECK();
    if ((st = launch_trueres(S, 0)) != GM_OK) return st;
    Scalars h;
    GM_HIP(hipMemcpyAsync(&h, S.w->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S.hs));
    GM_HIP(hipStreamSynchronize(S.hs));
    const double bb = h.bb > 0 ? h.bb : 1.0;
    *relres = sqrt(h.g0 / bb);
    return GM_OK;
}

gm_status gm_release_workspaces(void) {
    release_workspaces();
    release_red_scratch();
    return GM_OK;
}

}  // extern "C"

extern "C" {

gm_status gm_autotune(const BandedMatrix* L, int precision, int reps, gm_tune_result* out, gm_stream_t s) {
    Solver S{};
    gm_status st = get_plan(L, &S.plan);
    if (st != GM_OK) return st;
    if (precision < GM_PREC_FP32 || precision > GM_PREC_BF16) return fail(GM_ERR_INVALID, "autotune: precision");
    S.A = L;
    S.N = L->rows;
    S.hs = (hipStream_t)s;
    S.tiled = tiled_ok(S.plan, L->a, L->apitch);
    if (!S.tiled || S.N == 0) return fail(GM_ERR_UNSUPPORTED, "autotune: needs the 5x5 Levin band structure");
    const bool mixed = precision != GM_PREC_FP32;
    S.pm = mixed ? kScaled : kJacobi;
    S.fmt = precision == GM_PREC_FP16 ? kF16 : (precision == GM_PREC_BF16 ? kBF16 : kF32);
    if ((st = dev_info(&S.dv)) != GM_OK) return st;
    if ((st = red_scratch(s, &S.rs)) != GM_OK) return st;
    const bool f32fused = !mixed && fused32_ok(S.plan.W, S.N);
    if ((st = ws_get(s, S.N, S.plan.maxoff, mixed || f32fused, &S.w)) != GM_OK) return st;
    Workspace* w = S.w;
    S.g_elem = elementwise_grid(S.dv, S.N, 4);
    hipLaunchKernelGGL(k_fill, dim3(S.g_elem), dim3(kBlock), 0, S.hs, w->r, S.N, 1.0f);
    hipLaunchKernelGGL(k_fill, dim3(S.g_elem), dim3(kBlock), 0, S.hs, w->p0, S.N, 0.5f);
    hipLaunchKernelGGL(k_fill, dim3(S.g_elem), dim3(kBlock), 0, S.hs, w->p1, S.N, 0.5f);
    if (mixed) {
        S.sym = sym16_ok(S.plan.W, S.N);
        launch_scale16(S);
    } else {
        hipLaunchKernelGGL(k_diag_inv, dim3(S.g_elem), dim3(kBlock), 0, S.hs,
                           L->a + (size_t)S.plan.args.center * L->apitch, w->dinv, S.N);
    }
    GM_LAUNCH_CHECK();
    const int pcs[] = {1, 2, 3, 4, 6, 8};
    hipEvent_t e0, e1;
    GM_HIP(hipEventCreate(&e0));
    GM_HIP(hipEventCreate(&e1));
    TuneVal best{0, 0, 1e30, 0};
    if (reps < 1) reps = 5;
    char line[160];
    g_tune_report = "kernel shape R TY wg_per_cu us\n";
    g_tune_spmv = !mixed;
    for (int sh = 0; sh < kNumShapes; ++sh) {
        if (!step_shape_ok(sh, S.fmt, S.plan.W, S.N)) continue;
        for (int walk = 0; walk < (S.sym ? 2 : 1); ++walk) {
            for (int pc : pcs) {
                set_step_shape(S, sh, pc);
                S.walk = walk;
                double tot = 0;
                for (int k = 0; k < reps + 1; ++k) {
                    hipLaunchKernelGGL(k_tune_reset, dim3(1), dim3(64), 0, S.hs, w->sc);
                    GM_HIP(hipEventRecord(e0, S.hs));
                    if ((st = launch_step(S)) != GM_OK) return st;
                    GM_HIP(hipEventRecord(e1, S.hs));
                    GM_HIP(hipEventSynchronize(e1));
                    float ms = 0;
                    GM_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (k > 0) tot += ms;  // first launch warms caches
                }
                const double us = tot * 1e3 / reps;
                snprintf(line, sizeof(line), "cgstep %d %d %d %d %.2f%s\n", sh, kShapeR[sh], kShapeTY[sh], pc, us,
                         walk ? " column-walk" : "");
                g_tune_report += line;
                if (us < best.us) best = TuneVal{sh, pc, us, 0, walk};
            }
        }
    }
    // the update kernel's grid: fewer workgroups = fewer device-scope
    // fence + atomic arrivals in its reduction
    // one step so the update kernels below find a hand-off to run
    set_step_shape(S, best.shape, best.per_cu);
    S.walk = best.walk;
    hipLaunchKernelGGL(k_tune_reset, dim3(1), dim3(64), 0, S.hs, w->sc);
    if ((st = launch_step(S)) != GM_OK) return st;
    S.x = w->d;
    S.vec = true;
    double best_upd = 1e30;
    for (int pc : {1, 2, 4, 8, 16}) {
        S.g_upd = grid_for_tiles(S.dv, (S.N + 4L * kBlock - 1) / (4L * kBlock), pc);
        double tot = 0;
        for (int k = 0; k < reps + 1; ++k) {
            GM_HIP(hipEventRecord(e0, S.hs));
            if ((st = launch_update(S)) != GM_OK) return st;
            GM_HIP(hipEventRecord(e1, S.hs));
            GM_HIP(hipEventSynchronize(e1));
            float ms = 0;
            GM_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (k > 0) tot += ms;
        }
        const double us = tot * 1e3 / reps;
        snprintf(line, sizeof(line), "cgupdate - - - %d %.2f (grid %d)\n", pc, us, S.g_upd);
        g_tune_report += line;
        if (us < best_upd) {
            best_upd = us;
            best.upd_pc = pc;
        }
    }
    // the fused iteration in timing mode (fixed alpha and beta): mode 2 makes
    // its own and a deferred update of x, mode 5 defers its own; a solve
    // alternates the two, so their mean is the time per iteration
    double fused_us = 0, fused_upd_us = 0, fused_def_us = 0;
    auto time_fused = [&](int mode, double& us) -> gm_status {
        double tot = 0;
        for (int k = 0; k < reps + 1; ++k) {
            hipLaunchKernelGGL(k_tune_reset, dim3(1), dim3(64), 0, S.hs, w->sc);
            GM_HIP(hipEventRecord(e0, S.hs));
            gm_status r = launch_fused(S, mode);
            if (r != GM_OK) return r;
            GM_HIP(hipEventRecord(e1, S.hs));
            GM_HIP(hipEventSynchronize(e1));
            float ms = 0;
            GM_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (k > 0) tot += ms;
        }
        us = tot * 1e3 / reps;
        return GM_OK;
    };
    if (S.sym && S.N < kFusedMaxN) {
        // the fused iteration (timing mode: fixed alpha and beta; 4-row tiles,
        // column walk) against the best step + update, whichever step form won
        if ((st = ws_fused(w, S.N, S.plan.W, S.fmt)) != GM_OK) return st;
        launch_uedge(S);
        hipLaunchKernelGGL(k_fill, dim3(S.g_elem), dim3(kBlock), 0, S.hs, w->Ap, S.N, 0.25f);
        const int step_pc = best.per_cu;
        int fused_pc = 0;
        S.walk = 2;
        for (int pc : {2, 3}) {
            set_step_shape(S, kS2R4Y4, pc);
            double uf = 0, ud = 0;
            if ((st = time_fused(2, uf)) != GM_OK || (st = time_fused(5, ud)) != GM_OK) return st;
            const double us = 0.5 * (uf + ud);
            snprintf(line, sizeof(line),
                     "cgfused %d %d %d %d %.2f column-walk, mean of x updated %.2f / deferred %.2f (step + update %.2f)\n",
                     (int)kS2R4Y4, 4, 4, pc, us, uf, ud, best.us + best_upd);
            g_tune_report += line;
            if (fused_pc == 0 || us < fused_us) {
                fused_us = us;
                fused_upd_us = uf;
                fused_def_us = ud;
                fused_pc = pc;
            }
        }
        if (fused_us < best.us + best_upd) {
            best.shape = kS2R4Y4;
            best.walk = 2;
            best.per_cu = fused_pc;
        } else {
            best.per_cu = step_pc;
        }
        S.walk = best.walk;
        set_step_shape(S, best.shape, best.per_cu);
    }
    if (f32fused) {
        // the fused FP32 iteration on the symmetric scaled form (timing mode)
        // against the best Jacobi step + update
        launch_scale16(S);
        if ((st = ws_fused(w, S.N, S.plan.W, S.fmt)) != GM_OK) return st;
        launch_uedge(S);
        hipLaunchKernelGGL(k_fill, dim3(S.g_elem), dim3(kBlock), 0, S.hs, w->Ap, S.N, 0.25f);
        GM_LAUNCH_CHECK();
        const TuneVal step_best = best;
        int fused_pc = 0;
        S.walk = 2;
        for (int pc : {2, 3}) {
            set_step_shape(S, kS2R4Y4, pc);
            double uf = 0, ud = 0;
            if ((st = time_fused(2, uf)) != GM_OK || (st = time_fused(5, ud)) != GM_OK) return st;
            const double us = 0.5 * (uf + ud);
            snprintf(line, sizeof(line),
                     "cgfused %d %d %d %d %.2f FP32 symmetric column-walk, mean of x updated %.2f / deferred %.2f "
                     "(step + update %.2f)\n",
                     (int)kS2R4Y4, 2, 4, pc, us, uf, ud, step_best.us + best_upd);
            g_tune_report += line;
            if (fused_pc == 0 || us < fused_us) {
                fused_us = us;
                fused_upd_us = uf;
                fused_def_us = ud;
                fused_pc = pc;
            }
        }
        if (fused_us < step_best.us + best_upd) {
            best.shape = kS2R4Y4;
            best.walk = 2;
            best.per_cu = fused_pc;
        }
        S.walk = best.walk;
        set_step_shape(S, best.shape, best.per_cu);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune[TuneKey{S.fmt, S.plan.W, (long)S.N}] = best;
    }
    if (out) {
        out->update_wg_per_cu = best.upd_pc;
        out->update_us = best_upd;
        out->rows_per_thread = kShapeR[best.shape];
        out->tile_rows = kShapeTY[best.shape];
        out->wg_per_cu = best.per_cu;
        out->step_us = best.us;
        out->step_walk = best.walk;
        out->fused_us = fused_us;
        out->fused_update_us = fused_upd_us;
        out->fused_defer_us = fused_def_us;
        out->spmv_rows_per_thread = 0;
        out->spmv_tile_rows = 0;
        out->spmv_direct = 0;
        out->spmv_wg_per_cu = 0;
        out->spmv_us = 0;
    }
    if (!mixed) {
        // the plain FP32 product y = A x (gm_bm_ax) on the same matrix
        int sh = 0, pc = 0;
        double us = 0;
        if ((st = spmv_autotune(L->a, L->apitch, kF32, S.plan.W, S.N, w->Ap, w->r, reps, s, &sh, &pc, &us)) != GM_OK)
            return st;
        if (out) {
            spmv_shape_info(sh, &out->spmv_rows_per_thread, &out->spmv_tile_rows, &out->spmv_direct);
            out->spmv_wg_per_cu = pc;
            out->spmv_us = us;
        }
    }
    return GM_OK;
}

const char* gm_autotune_report(void) {
    static thread_local std::string full;
    full = g_tune_report + (g_tune_spmv ? spmv_report() : "");
    return full.c_str();
}

gm_status gm_spmv_tune_set(int W, int N, int rows_per_thread, int tile_rows, int direct, int wg_per_cu) {
    return spmv_tune_set(W, N, rows_per_thread, tile_rows, direct, wg_per_cu);
}

gm_status gm_tune_set_step(int W, int N, int precision, int rows_per_thread, int tile_rows, int wg_per_cu,
                           int column_walk) {
    const int fmt = precision == GM_PREC_FP16 ? kF16 : (precision == GM_PREC_BF16 ? kBF16 : kF32);
    int sh = -1;
    for (int k = 0; k < kNumShapes; ++k)
        if (kShapeR[k] == rows_per_thread && kShapeTY[k] == tile_rows) sh = k;
    if (sh < 0 || !step_shape_ok(sh, fmt, W, N) || wg_per_cu < 1 || wg_per_cu > 16)
        return fail(GM_ERR_INVALID, "gm_tune_set: unsupported shape");
    if (column_walk < 0 || column_walk > 2)
        return fail(GM_ERR_INVALID, "gm_tune_set_step: column_walk must be 0, 1 or 2");
    if (column_walk == 1 && (fmt == kF32 || !sym16_ok(W, N)))
        return fail(GM_ERR_INVALID, "gm_tune_set_step: the column walk is a form of the symmetric 16-bit step");
    if (column_walk == 2 && (fmt == kF32 ? !fused32_ok(W, N) : !sym16_ok(W, N)))
        return fail(GM_ERR_INVALID, "gm_tune_set_step: the fused iteration needs the symmetric layout (W %% 4 == 0)");
    if (column_walk == 2 && sh != kS2R4Y4)
        return fail(GM_ERR_INVALID, "gm_tune_set_step: the fused iteration walks 4-row tiles");
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[TuneKey{fmt, W, (long)N}] = TuneVal{sh, wg_per_cu, 0, 0, column_walk};
    return GM_OK;
}

gm_status gm_tune_set(int W, int N, int precision, int rows_per_thread, int tile_rows, int wg_per_cu) {
    return gm_tune_set_step(W, N, precision, rows_per_thread, tile_rows, wg_per_cu, 0);
}

}  // extern "C"
