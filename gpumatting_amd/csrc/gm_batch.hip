// gm_batch.hip -- many images of uneven size, several right-hand sides each,
// solved together (no reference counterpart).
//
// Each image i has one Levin Laplacian L_i (25 FP32 bands, no data term) and
// nrhs independent constraint sets j, i.e. nrhs systems
//     (L_i + diag(lam_ij)) x_ij = lam_ij * t_ij .
// The systems of one image share L_i's off-diagonal bands, so one pass over a
// tile loads those 24 bands into registers once and applies them to every
// right-hand side; each system brings its own FP32 diagonal d_ij = L_ii +
// lam_ij.  Jacobi-PCG per system, with the single-system kernels' sign and
// reduction conventions.
//
// One launch per phase covers all images: a global tile list (tile -> image)
// is walked by a persistent, XCD-aware grid.  Reductions are per image: every
// tile writes its partial sums, and the last tile of an image to arrive sums
// that image's tiles in index order and finalises its systems' scalars
// (deterministic).  Systems that have converged are skipped; an image whose
// systems are all done costs nothing.  Iterations are captured in a hipGraph
// (kernels only, no memset nodes) and replayed.
//
// Mixed precision: the off-diagonal bands are also kept in FP16 (unscaled;
// |L_ij| <= 9), the per-system diagonals stay FP32, and the residual is
// replaced by the FP64-accumulated true residual (FP32 L) every 50 iterations.
#include <string.h>

#include <vector>

#include "gm_internal.h"

namespace gm {

gm_status levin_assemble_raw(float* A, size_t apitch, float* b, float* dinv, const float* im, int nch, size_t pitch,
                             const float* lam, const float* tgt, const unsigned char* scr, float gamma, float eps,
                             int mode, int W, int H, gm_stream_t s);

constexpr int kMaxRhs = 8;
constexpr int kBR = 4;  // rows per thread of the batched kernels

struct BImg {
    int W, N, ntiles, tile0, sys0, nsys;
    const float* L;
    size_t pitch;
    const unsigned short* L16;
    size_t pitch16;
};

struct BSys {
    float *x, *r, *p0, *p1, *Ap, *d, *dinv, *b;
};

enum { kRun = 0, kConverged = 1, kBreakdown = 2, kNonFinite = 3, kMaxIt = 4 };

__device__ __forceinline__ bool b_stopped(const Scalars* s) { return s->done != kRun || s->verify != 0; }

// Sum part[(tile0 + k) * stride + off] over k < ntiles in index order; every
// thread gets the total, and every workgroup computes the same bits.
__device__ __forceinline__ double image_fold(const double* __restrict__ part, int tile0, int ntiles, int stride,
                                             int off) {
    __shared__ double fr[kWaves];
    __shared__ double ft;
    double s = 0;
    for (int k = threadIdx.x; k < ntiles; k += kBlock) s += part[(size_t)(tile0 + k) * stride + off];
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

// The batched iteration uses the same fence-free hand-off as the single
// system's CG kernels (see Scalars), per system: the step kernel folds the
// update kernel's per-tile partials of the image (once per image a workgroup
// visits) and only the workgroup holding the image's first tile writes the
// system's `hand`; the update kernel reads only `hand` and only that same
// workgroup publishes the system's Scalars.

// Step-kernel view of one system on entering an image: fold pending update
// partials and run the convergence test (every thread; results to chand).
__device__ __forceinline__ void b_step_enter(const BImg& I, const Scalars* sc, const double* __restrict__ part2,
                                             int verify, Hand* chand) {
    __syncthreads();  // chand of the previous image is no longer read
    for (int j = 0; j < I.nsys; ++j) {
        const Scalars* s = &sc[I.sys0 + j];
        Hand h{s->rz, s->rz_prev, s->rr, s->beta, s->iters, kHandRun, kRun, 0, I.ntiles};
        if (s->done != kRun || s->verify) {
            h.mode = kHandIdle;
        } else if (s->pend) {
            const double rz = image_fold(part2, I.tile0, I.ntiles, 2 * kMaxRhs, 2 * j);
            const double rr = image_fold(part2, I.tile0, I.ntiles, 2 * kMaxRhs, 2 * j + 1);
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

// K1: p = -D^-1 r + beta p_old staged per system, Ap = (L_off + diag d) p,
// per-tile <p, Ap> partials.
template <int FMT>
__global__ __launch_bounds__(kBlock) void kb_cgstep(const BImg* __restrict__ imgs, const BSys* __restrict__ sys,
                                                    Scalars* sc, const int* __restrict__ tile_img, int ntot,
                                                    double* __restrict__ part1, const double* __restrict__ part2,
                                                    int verify) {
    constexpr int R = kBR;
    __shared__ __attribute__((aligned(16))) float lds[Tile<R>::LDS_FLOATS];
    __shared__ Hand chand[kMaxRhs];
    __shared__ double part[kMaxRhs];
    const TileRange tr = xcd_tiles(ntot);
    int cimg = -1;
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int im = tile_img[t];
        const BImg I = imgs[im];
        if (im != cimg) {
            b_step_enter(I, sc, part2, verify, chand);
            cimg = im;
        }
        if (t == I.tile0 && threadIdx.x == 0)
            for (int j = 0; j < I.nsys; ++j) sc[I.sys0 + j].hand = chand[j];
        unsigned mask = 0;
        for (int j = 0; j < I.nsys; ++j)
            if (chand[j].mode == kHandRun) mask |= 1u << j;
        if (mask == 0) continue;
        const int i0 = (t - I.tile0) * Tile<R>::ROWS;
        const int ib = i0 + R * threadIdx.x;
        Bands<R, FMT> av;
        if (ib < I.N) {
            if (FMT == kF32) load_bands<R, FMT, kDiagCaller>(I.L, I.pitch, ib, av);
            else load_bands<R, FMT, kDiagCaller>(I.L16, I.pitch16, ib, av);
        }
        for (int j = 0; j < I.nsys; ++j) {
            if (!((mask >> j) & 1u)) {
                if (threadIdx.x == 0) part[j] = 0;
                continue;
            }
            const BSys S = sys[I.sys0 + j];
            const float beta = chand[j].beta;
            const int par = chand[j].iters & 1;
            const float* __restrict__ pold = par ? S.p1 : S.p0;
            float* __restrict__ pnew = par ? S.p0 : S.p1;
            __syncthreads();
            if (beta != 0.f)
                stage5<R>(lds, i0, I.W, I.N, [&](int g) { return fmaf(beta, pold[g], -(S.dinv[g] * S.r[g])); });
            else
                stage5<R>(lds, i0, I.W, I.N, [&](int g) { return -(S.dinv[g] * S.r[g]); });
            __syncthreads();
            double v = 0;
            if (ib < I.N) {
                float y[R], pc[R], dd[R];
                apply_bands<R, FMT, kDiagCaller>(av, lds + R * threadIdx.x, y);
                load_rows<R>(S.d, ib, I.N, dd);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    pc[q] = lds[2 * Tile<R>::SEG + R * threadIdx.x + 2 + q];
                    y[q] = fmaf(dd[q], pc[q], y[q]);
                }
                store_rows<R>(pnew, ib, I.N, pc);
                store_rows<R>(S.Ap, ib, I.N, y);
#pragma unroll
                for (int q = 0; q < R; ++q)
                    if (ib + q < I.N) v += (double)(pc[q] * y[q]);
            }
            const double bs = block_sum(v);
            if (threadIdx.x == 0) part[j] = bs;
        }
        __syncthreads();
        if (threadIdx.x < I.nsys) part1[(size_t)t * kMaxRhs + threadIdx.x] = part[threadIdx.x];
        __syncthreads();  // part is rewritten by the next tile
    }
}

enum { kBRun = 0, kBIdle = 10, kBStop = 11 };  // update-kernel outcome per system (else kBreakdown / kNonFinite)

// K2: x += alpha p, r += alpha Ap, per-tile <r, D^-1 r>, ||r||^2 partials.
__global__ __launch_bounds__(kBlock) void kb_cgupdate(const BImg* __restrict__ imgs, const BSys* __restrict__ sys,
                                                      Scalars* sc, const int* __restrict__ tile_img, int ntot,
                                                      const double* __restrict__ part1, double* __restrict__ part2) {
    constexpr int R = kBR;
    __shared__ Hand chand[kMaxRhs];
    __shared__ int cst[kMaxRhs];
    __shared__ float calpha[kMaxRhs];
    __shared__ double cpap[kMaxRhs];
    __shared__ double part[kMaxRhs][2];
    const TileRange tr = xcd_tiles(ntot);
    int cimg = -1;
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int im = tile_img[t];
        const BImg I = imgs[im];
        if (im != cimg) {
            __syncthreads();
            for (int j = 0; j < I.nsys; ++j) {
                const Hand h = sc[I.sys0 + j].hand;
                int st = kBRun;
                float a = 0.f;
                double pAp = 0;
                if (h.mode == kHandIdle) {
                    st = kBIdle;
                } else if (h.mode == kHandStop) {
                    st = kBStop;
                } else {
                    pAp = image_fold(part1, I.tile0, I.ntiles, kMaxRhs, j);
                    if (!isfinite(pAp)) st = kNonFinite;
                    else if (!(pAp > 0)) st = kBreakdown;
                    else a = (float)(h.rz / pAp);
                }
                if (threadIdx.x == 0) {
                    chand[j] = h;
                    cst[j] = st;
                    calpha[j] = a;
                    cpap[j] = pAp;
                }
            }
            __syncthreads();
            cimg = im;
        }
        if (t == I.tile0 && threadIdx.x == 0)
            for (int j = 0; j < I.nsys; ++j) {
                Scalars* s = &sc[I.sys0 + j];
                const Hand& h = chand[j];
                if (cst[j] == kBStop) {
                    publish(s, h, h.done, h.verify, h.iters, 0, 0);
                } else if (cst[j] == kBRun) {
                    publish(s, h, kRun, 0, h.iters + 1, 1, I.ntiles);
                    s->pAp = cpap[j];
                    s->alpha = calpha[j];
                } else if (cst[j] != kBIdle) {
                    publish(s, h, cst[j], 0, h.iters, 0, 0);
                    s->pAp = cpap[j];
                }
            }
        unsigned mask = 0;
        for (int j = 0; j < I.nsys; ++j)
            if (cst[j] == kBRun) mask |= 1u << j;
        if (mask == 0) continue;
        const int i0 = (t - I.tile0) * Tile<R>::ROWS;
        for (int j = 0; j < I.nsys; ++j) {
            if (!((mask >> j) & 1u)) {
                if (threadIdx.x == 0) part[j][0] = part[j][1] = 0;
                continue;
            }
            const BSys S = sys[I.sys0 + j];
            const float alpha = calpha[j];
            const float* __restrict__ p = (chand[j].iters & 1) ? S.p0 : S.p1;
            double v0 = 0, v1 = 0;
            for (int q = 0; q < R; ++q) {
                const int i = i0 + q * kBlock + threadIdx.x;  // coalesced across the workgroup
                if (i >= I.N) break;
                S.x[i] = fmaf(alpha, p[i], S.x[i]);
                const float rn = fmaf(alpha, S.Ap[i], S.r[i]);
                S.r[i] = rn;
                v0 += (double)(rn * (S.dinv[i] * rn));
                v1 += (double)(rn * rn);
            }
            const double b0 = block_sum(v0);
            const double b1 = block_sum(v1);
            if (threadIdx.x == 0) {
                part[j][0] = b0;
                part[j][1] = b1;
            }
        }
        __syncthreads();
        if (threadIdx.x < 2 * I.nsys)
            part2[(size_t)t * 2 * kMaxRhs + threadIdx.x] = part[threadIdx.x >> 1][threadIdx.x & 1];
        __syncthreads();  // part is rewritten by the next tile
    }
}

// Fold pending update partials into the Scalars (one workgroup per image;
// launched before the host reads the final state).
__global__ __launch_bounds__(kBlock) void kb_fold(const BImg* __restrict__ imgs, Scalars* sc,
                                                  const double* __restrict__ part2, int verify) {
    const BImg I = imgs[blockIdx.x];
    for (int j = 0; j < I.nsys; ++j) {
        Scalars* s = &sc[I.sys0 + j];
        if (s->done != kRun || s->verify || !s->pend) continue;
        const double rz = image_fold(part2, I.tile0, I.ntiles, 2 * kMaxRhs, 2 * j);
        const double rr = image_fold(part2, I.tile0, I.ntiles, 2 * kMaxRhs, 2 * j + 1);
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

// FP64-accumulated true residual t = (L_off + diag d) x - b per system, per-tile
// partials (<t, D^-1 t>, ||t||^2, drift ||t - r||^2); kb_trueres_fin folds
// them.  op 0: measure only (g0); 1: replace r for running systems; 2: start
// (all systems, beta = 0); 3: replace r for systems whose recursive residual
// asked for verification.
__global__ __launch_bounds__(kBlock) void kb_trueres(const BImg* __restrict__ imgs, const BSys* __restrict__ sys,
                                                     const Scalars* sc, const int* __restrict__ tile_img, int ntot,
                                                     int op, double* __restrict__ part3) {
    constexpr int R = kBR;
    constexpr int SEG = Tile<R>::SEG;
    __shared__ __attribute__((aligned(16))) float lds[Tile<R>::LDS_FLOATS];
    __shared__ double part[kMaxRhs][3];
    const TileRange tr = xcd_tiles(ntot);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int im = tile_img[t];
        const BImg I = imgs[im];
        unsigned mask = 0;
        for (int j = 0; j < I.nsys; ++j)
            if (tr_selected(&sc[I.sys0 + j], op)) mask |= 1u << j;
        if (mask == 0) continue;
        const int i0 = (t - I.tile0) * Tile<R>::ROWS;
        const int ib = i0 + R * threadIdx.x;
        Bands<R, kF32> av;
        if (ib < I.N) load_bands<R, kF32, kDiagCaller>(I.L, I.pitch, ib, av);
        for (int j = 0; j < I.nsys; ++j) {
            if (!((mask >> j) & 1u)) {
                if (threadIdx.x == 0) part[j][0] = part[j][1] = part[j][2] = 0;
                continue;
            }
            const BSys S = sys[I.sys0 + j];
            // periodic replacement of a running system: measure the drift
            const bool drift = op == 1 && sc[I.sys0 + j].verify == 0;
            __syncthreads();
            stage5<R>(lds, i0, I.W, I.N, [&](int g) { return S.x[g]; });
            __syncthreads();
            double v0 = 0, v1 = 0, v2 = 0;
            if (ib < I.N) {
                const float* xs = lds + R * threadIdx.x;
                double acc[R];
#pragma unroll
                for (int q = 0; q < R; ++q) acc[q] = 0.0;
#pragma unroll
                for (int k = 0; k < 25; ++k) {
                    if (k == 12) continue;
                    const int s5 = k / 5, dx = k % 5;
#pragma unroll
                    for (int q = 0; q < R; ++q)
                        acc[q] = fma((double)av.v[k].get(q), (double)xs[s5 * SEG + q + dx], acc[q]);
                }
                float dd[R], bb[R], di[R], out[R], ro[R];
                load_rows<R>(S.d, ib, I.N, dd);
                load_rows<R>(S.b, ib, I.N, bb);
                load_rows<R>(S.dinv, ib, I.N, di);
                if (drift) load_rows<R>(S.r, ib, I.N, ro);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const double tq = fma((double)dd[q], (double)xs[2 * SEG + q + 2], acc[q]) - (double)bb[q];
                    out[q] = (float)tq;
                    if (ib + q < I.N) {
                        v1 += tq * tq;
                        v0 += (double)(out[q] * (di[q] * out[q]));
                        if (drift) v2 += (double)(out[q] - ro[q]) * (double)(out[q] - ro[q]);
                    }
                }
                if (op != 0) store_rows<R>(S.r, ib, I.N, out);
            }
            const double b0 = block_sum(v0);
            const double b1 = block_sum(v1);
            const double b2 = block_sum(v2);
            if (threadIdx.x == 0) {
                part[j][0] = b0;
                part[j][1] = b1;
                part[j][2] = b2;
            }
        }
        __syncthreads();
        if (threadIdx.x < 3 * I.nsys)
            part3[(size_t)t * 3 * kMaxRhs + threadIdx.x] = part[threadIdx.x / 3][threadIdx.x % 3];
        __syncthreads();  // part is rewritten by the next tile
    }
}

// Finish a true-residual pass: one workgroup per image folds its tiles'
// partials and updates the selected systems' scalars.
__global__ __launch_bounds__(kBlock) void kb_trueres_fin(const BImg* __restrict__ imgs, Scalars* sc,
                                                         const double* __restrict__ part3, int op) {
    const BImg I = imgs[blockIdx.x];
    for (int j = 0; j < I.nsys; ++j) {
        Scalars* s = &sc[I.sys0 + j];
        if (!tr_selected(s, op)) continue;
        const double rz = image_fold(part3, I.tile0, I.ntiles, 3 * kMaxRhs, 3 * j);
        const double rr = image_fold(part3, I.tile0, I.ntiles, 3 * kMaxRhs, 3 * j + 1);
        const double d2 = image_fold(part3, I.tile0, I.ntiles, 3 * kMaxRhs, 3 * j + 2);
        if (threadIdx.x != 0) continue;
        s->g0 = rr;
        if (op == 0) continue;
        if (!isfinite(rz) || !isfinite(rr)) {
            s->done = kNonFinite;
            continue;
        }
        if (op == 2) {
            s->beta = 0.f;
        } else if (s->verify) {
            // drifted recursion: restart from the true residual
            s->beta = 0.f;
            s->replacements += 1;
            if (s->g1 > 0 && rr > 0.25 * s->g1) s->stall += 1;
            else s->stall = 0;
            s->g1 = rr;
        } else {
            // periodic: keep the direction unless the recursion drifted; the
            // <r,z> of the last step is rz while its update's partials are
            // pending (superseded by this residual), else rz_prev
            const double rzp = s->pend ? s->rz : s->rz_prev;
            s->beta = d2 > kDrift2 * rr ? 0.f : (rzp > 0 ? (float)(rz / rzp) : 0.f);
            s->replacements += 1;
        }
        s->rz_prev = s->pend ? s->rz : s->rz_prev;
        s->rz = rz;
        s->rr = rr;
        s->verify = 0;
        s->pend = 0;
        if (s->tol2 >= 0 && rr <= s->tol2) s->done = kConverged;
        else if (s->iters >= s->maxit || s->stall >= 3) s->done = kMaxIt;
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

__global__ __launch_bounds__(kBlock) void kb_to16(const float* __restrict__ L, size_t pitch, int N,
                                                  unsigned short* __restrict__ L16, size_t pitch16) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock)
#pragma unroll
        for (int k = 0; k < 25; ++k)
            if (k != 12) L16[(size_t)k * pitch16 + i] = f32_to_f16_rne(L[(size_t)k * pitch + i]);
}

__global__ void kb_init(Scalars* sc, int nsys, int maxit, double tol) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsys) return;
    Scalars z = {};
    z.maxit = maxit;
    z.tol2 = -1.0;
    z.hand.mode = kHandIdle;
    sc[j] = z;
    (void)tol;
}

// per system <b,b> -> tol2 (one workgroup per system)
__global__ __launch_bounds__(kBlock) void kb_bnorm(const BSys* __restrict__ sys, const int* __restrict__ sysN,
                                                   Scalars* sc, double tol) {
    const int j = blockIdx.x;
    const BSys S = sys[j];
    const int N = sysN[j];
    double v = 0;
    for (int i = threadIdx.x; i < N; i += kBlock) v += (double)S.b[i] * (double)S.b[i];
    const double bb = block_sum(v);
    if (threadIdx.x == 0) {
        sc[j].bb = bb;
        sc[j].tol2 = tol > 0 ? tol * tol * (bb > 0 ? bb : 1.0) : -1.0;
    }
}

}  // namespace gm

using namespace gm;

struct gm_batch_s {
    int nimg = 0, nrhs = 0, nsys = 0, precision = 0, ntiles = 0;
    std::vector<BImg> himg;
    std::vector<BSys> hsys;
    std::vector<int> hsysN;
    BImg* dimg = nullptr;
    BSys* dsys = nullptr;
    int* dsysN = nullptr;
    Scalars* dsc = nullptr;
    Scalars* hsc = nullptr;  // pinned, nsys
    int* dtile = nullptr;
    double* dpart = nullptr;  // per-tile partial sums: step, update (2), true residual (3)
    float* Lmem = nullptr;
    unsigned short* L16mem = nullptr;
    float* vmem = nullptr;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_iters = 0, graph_grid = 0, graph_verify = -1;
    hipEvent_t ev = nullptr;
    int per_cu = 2;
};

static void batch_free(gm_batch* b) {
    if (!b) return;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->ev) (void)hipEventDestroy(b->ev);
    for (void* p : {(void*)b->dimg, (void*)b->dsys, (void*)b->dsysN, (void*)b->dsc, (void*)b->dtile, (void*)b->dpart,
                    (void*)b->Lmem, (void*)b->L16mem, (void*)b->vmem})
        if (p) (void)hipFree(p);
    if (b->hsc) (void)hipHostFree(b->hsc);
    delete b;
}

extern "C" {

gm_status gm_batch_create(gm_batch** out, int nimages, const int* widths, const int* heights, int nrhs, int precision) {
    if (!out || nimages <= 0 || !widths || !heights || nrhs <= 0 || nrhs > kMaxRhs)
        return fail(GM_ERR_INVALID, "gm_batch_create: bad arguments (nrhs must be 1..%d)", kMaxRhs);
    if (precision != GM_PREC_FP32 && precision != GM_PREC_FP16)
        return fail(GM_ERR_UNSUPPORTED, "gm_batch_create: batched forms are FP32 and FP16");
    gm_batch* b = new gm_batch();
    b->nimg = nimages;
    b->nrhs = nrhs;
    b->nsys = nimages * nrhs;
    b->precision = precision;
    size_t Lfloats = 0, L16n = 0, vfloats = 0;
    int tiles = 0;
    std::vector<size_t> loff(nimages), l16off(nimages), voff(nimages);
    for (int i = 0; i < nimages; ++i) {
        const int W = widths[i], H = heights[i];
        if (W < 1 || H < 1 || (long long)W * H > (1LL << 30)) {
            batch_free(b);
            return fail(GM_ERR_INVALID, "gm_batch_create: image %d is %dx%d", i, W, H);
        }
        const int N = W * H;
        const size_t pitch = (size_t)(N + 63) / 64 * 64;
        BImg I{};
        I.W = W;
        I.N = N;
        I.ntiles = (N + Tile<kBR>::ROWS - 1) / Tile<kBR>::ROWS;
        I.tile0 = tiles;
        I.sys0 = i * nrhs;
        I.nsys = nrhs;
        I.pitch = pitch;
        I.pitch16 = pitch;
        tiles += I.ntiles;
        loff[i] = Lfloats;
        Lfloats += 25 * pitch;
        l16off[i] = L16n;
        if (precision == GM_PREC_FP16) L16n += 25 * pitch;
        voff[i] = vfloats;
        vfloats += (size_t)nrhs * 8 * pitch;
        b->himg.push_back(I);
    }
    b->ntiles = tiles;
    hipError_t e = hipSuccess;
#define BTRY(x)                                                        \
    do {                                                               \
        e = (x);                                                       \
        if (e != hipSuccess) {                                         \
            batch_free(b);                                             \
            return hip_fail(e, #x, __FILE__, __LINE__);                \
        }                                                              \
    } while (0)
    BTRY(hipMalloc(&b->Lmem, sizeof(float) * Lfloats + 256));
    BTRY(hipMemset(b->Lmem, 0, sizeof(float) * Lfloats + 256));
    if (L16n) {
        BTRY(hipMalloc(&b->L16mem, sizeof(unsigned short) * L16n + 256));
        BTRY(hipMemset(b->L16mem, 0, sizeof(unsigned short) * L16n + 256));
    }
    BTRY(hipMalloc(&b->vmem, sizeof(float) * vfloats + 256));
    BTRY(hipMemset(b->vmem, 0, sizeof(float) * vfloats + 256));
    for (int i = 0; i < nimages; ++i) {
        BImg& I = b->himg[i];
        I.L = b->Lmem + loff[i];
        I.L16 = L16n ? b->L16mem + l16off[i] : nullptr;
        for (int j = 0; j < nrhs; ++j) {
            float* base = b->vmem + voff[i] + (size_t)j * 8 * I.pitch;
            BSys S;
            S.x = base;
            S.r = base + 1 * I.pitch;
            S.p0 = base + 2 * I.pitch;
            S.p1 = base + 3 * I.pitch;
            S.Ap = base + 4 * I.pitch;
            S.d = base + 5 * I.pitch;
            S.dinv = base + 6 * I.pitch;
            S.b = base + 7 * I.pitch;
            b->hsys.push_back(S);
            b->hsysN.push_back(I.N);
        }
    }
    std::vector<int> tile_img(tiles);
    for (int i = 0; i < nimages; ++i)
        for (int k = 0; k < b->himg[i].ntiles; ++k) tile_img[b->himg[i].tile0 + k] = i;
    BTRY(hipMalloc(&b->dimg, sizeof(BImg) * nimages));
    BTRY(hipMalloc(&b->dsys, sizeof(BSys) * b->nsys));
    BTRY(hipMalloc(&b->dsysN, sizeof(int) * b->nsys));
    BTRY(hipMalloc(&b->dsc, sizeof(Scalars) * b->nsys));
    BTRY(hipMemset(b->dsc, 0, sizeof(Scalars) * b->nsys));
    BTRY(hipMalloc(&b->dtile, sizeof(int) * tiles));
    const size_t nparts = (size_t)tiles * kMaxRhs * 6;
    BTRY(hipMalloc(&b->dpart, sizeof(double) * nparts));
    BTRY(hipMemset(b->dpart, 0, sizeof(double) * nparts));
    BTRY(hipMemcpy(b->dimg, b->himg.data(), sizeof(BImg) * nimages, hipMemcpyHostToDevice));
    BTRY(hipMemcpy(b->dsys, b->hsys.data(), sizeof(BSys) * b->nsys, hipMemcpyHostToDevice));
    BTRY(hipMemcpy(b->dsysN, b->hsysN.data(), sizeof(int) * b->nsys, hipMemcpyHostToDevice));
    BTRY(hipMemcpy(b->dtile, tile_img.data(), sizeof(int) * tiles, hipMemcpyHostToDevice));
    BTRY(hipHostMalloc(&b->hsc, sizeof(Scalars) * b->nsys, hipHostMallocDefault));
    BTRY(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    BTRY(hipEventCreateWithFlags(&b->ev, hipEventDisableTiming));
#undef BTRY
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
    const int H = I.N / I.W;
    gm_status st = levin_assemble_raw(const_cast<float*>(I.L), I.pitch, nullptr, nullptr, im, nch, im_pitch, nullptr,
                                      nullptr, nullptr, 0.f, eps, GM_ASSEMBLY_ACCURATE, I.W, H, s);
    if (st != GM_OK) return st;
    if (b->precision == GM_PREC_FP16) {
        DevInfo d;
        if ((st = dev_info(&d)) != GM_OK) return st;
        hipLaunchKernelGGL(kb_to16, dim3(elementwise_grid(d, I.N, 1)), dim3(kBlock), 0, (hipStream_t)s, I.L, I.pitch,
                           I.N, const_cast<unsigned short*>(I.L16), I.pitch16);
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

static double* part1(gm_batch* b) { return b->dpart; }
static double* part2(gm_batch* b) { return b->dpart + (size_t)b->ntiles * kMaxRhs; }
static double* part3(gm_batch* b) { return b->dpart + (size_t)b->ntiles * kMaxRhs * 3; }

static gm_status batch_launch_iter(gm_batch* b, int grid, int verify, hipStream_t hs) {
    if (b->precision == GM_PREC_FP16)
        hipLaunchKernelGGL(kb_cgstep<kF16>, dim3(grid), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, b->dtile,
                           b->ntiles, part1(b), part2(b), verify);
    else
        hipLaunchKernelGGL(kb_cgstep<kF32>, dim3(grid), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, b->dtile,
                           b->ntiles, part1(b), part2(b), verify);
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_cgupdate, dim3(grid), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, b->dtile, b->ntiles,
                       part1(b), part2(b));
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status batch_trueres(gm_batch* b, int grid, int op, hipStream_t hs) {
    hipLaunchKernelGGL(kb_trueres, dim3(grid), dim3(kBlock), 0, hs, b->dimg, b->dsys, b->dsc, b->dtile, b->ntiles, op,
                       part3(b));
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_trueres_fin, dim3(b->nimg), dim3(kBlock), 0, hs, b->dimg, b->dsc, part3(b), op);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status batch_fold(gm_batch* b, int verify, hipStream_t hs) {
    hipLaunchKernelGGL(kb_fold, dim3(b->nimg), dim3(kBlock), 0, hs, b->dimg, b->dsc, part2(b), verify);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

// Capture `iters` iterations (K1, K2 pairs) into a graph on the batch's own
// stream (capture needs a non-null stream); replayed on the caller's stream.
static gm_status batch_graph(gm_batch* b, int iters, int grid, int verify) {
    if (b->gexec && b->graph_iters == iters && b->graph_grid == grid && b->graph_verify == verify) return GM_OK;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    GM_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    gm_status st = GM_OK;
    for (int k = 0; k < iters && st == GM_OK; ++k) st = batch_launch_iter(b, grid, verify, b->cap);
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
    b->graph_grid = grid;
    b->graph_verify = verify;
    return GM_OK;
}

gm_status gm_batch_solve(gm_batch* b, const gm_solve_opts* opts, gm_stream_t s) {
    if (!b) return fail(GM_ERR_INVALID, "gm_batch_solve: null batch");
    gm_solve_opts o;
    if (opts) o = *opts;
    else gm_solve_opts_default(&o);
    if (o.method != GM_METHOD_PCG) return fail(GM_ERR_UNSUPPORTED, "gm_batch_solve: batched solver is Jacobi-PCG");
    if (o.max_iter <= 0) return fail(GM_ERR_INVALID, "gm_batch_solve: max_iter");
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    hipStream_t hs = (hipStream_t)s;
    const int grid = grid_for_tiles(d, b->ntiles, b->per_cu);
    const int verify = (o.tol > 0 && o.verify) ? 1 : 0;
    const int chunk = o.check_every > 0 ? o.check_every : 10;
    const bool mixed = b->precision == GM_PREC_FP16;
    int refine = o.refine_every > 0 ? o.refine_every : (o.refine_every == 0 && mixed ? 50 : 0);
    if (refine > 0 && refine % chunk) refine = (refine / chunk + 1) * chunk;  // replacement at chunk boundaries

    hipLaunchKernelGGL(kb_init, dim3((b->nsys + 63) / 64), dim3(64), 0, hs, b->dsc, b->nsys, o.max_iter, o.tol);
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_bnorm, dim3(b->nsys), dim3(kBlock), 0, hs, b->dsys, b->dsysN, b->dsc, o.tol);
    GM_LAUNCH_CHECK();
    if ((st = batch_trueres(b, grid, 2, hs)) != GM_OK) return st;
    if ((st = batch_graph(b, chunk, grid, verify)) != GM_OK) return st;

    long enq = 0;
    bool finished = false;
    while (!finished) {
        GM_HIP(hipGraphLaunch(b->gexec, hs));
        enq += chunk;
        if (refine > 0 && enq % refine == 0) {
            if ((st = batch_trueres(b, grid, 1, hs)) != GM_OK) return st;
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
            if ((st = batch_trueres(b, grid, 3, hs)) != GM_OK) return st;
        }
    }
    // fold the last update's partials, check any recursive convergence it
    // shows, then the FP64 true residual of every system (g0) and a host copy
    if ((st = batch_fold(b, verify, hs)) != GM_OK) return st;
    if ((st = batch_trueres(b, grid, 3, hs)) != GM_OK) return st;
    if ((st = batch_trueres(b, grid, 0, hs)) != GM_OK) return st;
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
    if (!b || per_cu < 1 || per_cu > 16) return fail(GM_ERR_INVALID, "gm_batch_set_grid: per_cu 1..16");
    b->per_cu = per_cu;
    return GM_OK;
}

int gm_batch_tiles(gm_batch* b) { return b ? b->ntiles : 0; }

}  // extern "C"
