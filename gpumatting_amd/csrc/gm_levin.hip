// gm_levin.hip -- Levin matting Laplacian assembly, gather form.
//
// Replaces the reference's single-threaded host scatter (hostLevinLaplacian,
// Matting.cu:39-219) and its unfinished GPU scatter kernel (Matting.cu:240-382,
// which raced on L).  Each thread owns one row i of L (one pixel) and
// collects, for every interior 3x3 window containing the pixel (window centres
// in raster order), that window's 9 contributions -(1 + w_i.(x_j - mu))/9 into
// 25 register accumulators.  An entry L[i, band] therefore receives the same
// float additions in the same order as the reference's scatter loop, which is
// what makes GM_ASSEMBLY_REFERENCE bit-identical to it; no atomics, no races.
//
// GM_ASSEMBLY_ACCURATE keeps the same gather but evaluates each window's
// contribution from FP64 statistics: per window S' = -(C + eps/9 I)^-1 / 9
// and the mean mu = mh + ml (float head and tail); per pixel i and window w
// the whitened vector v = S' (x_i - mu) (double, rounded to float once) and
// c = -1/9 - v.ml, so that each other member j adds
// c + v.(x_j - mh) = -1/9 + v.(x_j - mu) -- one subtraction and one FMA per
// channel -- to L[i, j].  The operations (and their FMAs) are fixed, so the
// result is deterministic and restated exactly by the CPU oracle.
//
// Workgroup tile: 32 x 8 pixels (a wave's 64 pixels are 2 rows of 32, so
// each of its band stores writes two whole 128-byte lines).  LDS holds the
// 36 x 12 pixel halo and the statistics (mean, FP64 inverse covariance) of the
// 34 x 10 windows touching the tile; each window's statistics are computed
// once per workgroup.  ACCURATE keeps 30 KB of LDS and 92 VGPRs: 5 workgroups
// per CU.  Output per pixel: 25 diagonals (100 B), b, 1/L_ii.
#include "gm_internal.h"

namespace gm {

constexpr int kTX = 32, kTY = 8;              // pixels per workgroup
constexpr int kWX = kTX + 2, kWY = kTY + 2;   // windows touching the tile
constexpr int kPX = kTX + 4, kPY = kTY + 4;   // pixels those windows read
constexpr int kNW = kWX * kWY;
constexpr double kInv9 = 1.0 / 9.0;
constexpr float kInv9f = 1.0f / 9.0f;

// MODE: GM_ASSEMBLY_REFERENCE or GM_ASSEMBLY_ACCURATE
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_levin(const float* __restrict__ im, int nch, size_t pitch,
                                                  const float* __restrict__ lam, const float* __restrict__ tgt,
                                                  const unsigned char* __restrict__ scr, float gamma, float eps,
                                                  int W, int H, float* __restrict__ A, size_t apitch,
                                                  float* __restrict__ bvec, float* __restrict__ dinv) {
#pragma clang fp contract(off)
    __shared__ float px[3][kPY][kPX];
    // per-window statistics, structure of arrays:
    //   REFERENCE: ws[0..5] = adjugate d11 d12 d13 d22 d23 d33, ws[6] = det,
    //              ws[7..9] = float mean (held exactly as double)
    //   ACCURATE : ws[0..5] = -inverse/9: S'11 S'12 S'13 S'22 S'23 S'33
    __shared__ double ws[MODE == GM_ASSEMBLY_ACCURATE ? 6 : 10][kNW];
    __shared__ float wm[MODE == GM_ASSEMBLY_ACCURATE ? 6 : 1][kNW];  // ACCURATE: mean = head wm[0..2] + tail wm[3..5]
    __shared__ unsigned char wvalid[kNW];

    const int x0 = blockIdx.x * kTX, y0 = blockIdx.y * kTY;

    for (int e = threadIdx.x; e < kPX * kPY; e += kBlock) {
        const int ly = e / kPX, lx = e % kPX;
        const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
        float c0 = 0.f, c1 = 0.f, c2 = 0.f;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const float* p = im + ((size_t)gy * pitch + (size_t)gx) * (size_t)nch;
            c0 = p[0]; c1 = p[1]; c2 = p[2];
        }
        px[0][ly][lx] = c0; px[1][ly][lx] = c1; px[2][ly][lx] = c2;
    }
    __syncthreads();

    for (int e = threadIdx.x; e < kNW; e += kBlock) {
        const int wy = e / kWX, wx = e % kWX;
        const int cx = x0 - 1 + wx, cy = y0 - 1 + wy;
        const bool ok = cx >= 1 && cx <= W - 2 && cy >= 1 && cy <= H - 2;
        wvalid[e] = ok;
        if (!ok) continue;
        if constexpr (MODE == GM_ASSEMBLY_REFERENCE) {
            // Matting.cu:95-137: float mean, double sums of float products,
            // eps/9 in float, double determinant and adjugate.
            float mx = 0.f, my = 0.f, mz = 0.f;
            double c11 = 0, c12 = 0, c13 = 0, c22 = 0, c23 = 0, c33 = 0;
#pragma unroll
            for (int v1 = 0; v1 < 3; ++v1)
#pragma unroll
                for (int u1 = 0; u1 < 3; ++u1) {
                    const float x = px[0][wy + v1][wx + u1], y = px[1][wy + v1][wx + u1], z = px[2][wy + v1][wx + u1];
                    mx += x; my += y; mz += z;
                    float t;
                    t = x * x; c11 += (double)t;
                    t = x * y; c12 += (double)t;
                    t = x * z; c13 += (double)t;
                    t = y * y; c22 += (double)t;
                    t = y * z; c23 += (double)t;
                    t = z * z; c33 += (double)t;
                }
            mx /= 9.0f; my /= 9.0f; mz /= 9.0f;
            const float reg = eps / 9.0f;
            float t;
            t = mx * mx; c11 = c11 / 9.0 - (double)t + (double)reg;
            t = mx * my; c12 = c12 / 9.0 - (double)t;
            t = mx * mz; c13 = c13 / 9.0 - (double)t;
            t = my * my; c22 = c22 / 9.0 - (double)t + (double)reg;
            t = my * mz; c23 = c23 / 9.0 - (double)t;
            t = mz * mz; c33 = c33 / 9.0 - (double)t + (double)reg;
            ws[6][e] = c11 * c22 * c33 + c12 * c23 * c13 + c13 * c12 * c23 - c13 * c22 * c13 - c12 * c12 * c33 -
                       c11 * c23 * c23;
            ws[0][e] = c33 * c22 - c23 * c23;
            ws[1][e] = c13 * c23 - c33 * c12;
            ws[2][e] = c23 * c12 - c22 * c13;
            ws[3][e] = c33 * c11 - c13 * c13;
            ws[4][e] = c12 * c13 - c23 * c11;
            ws[5][e] = c22 * c11 - c12 * c12;
            ws[7][e] = (double)mx; ws[8][e] = (double)my; ws[9][e] = (double)mz;
        } else {
            // exact double sums (float products are exact in double), the
            // 1/9 scalings as multiplications, FMA-contracted adjugate and
            // determinant, one division
            double m0 = 0, m1 = 0, m2 = 0, s00 = 0, s01 = 0, s02 = 0, s11 = 0, s12 = 0, s22 = 0;
#pragma unroll
            for (int v1 = 0; v1 < 3; ++v1)
#pragma unroll
                for (int u1 = 0; u1 < 3; ++u1) {
                    const double x = px[0][wy + v1][wx + u1], y = px[1][wy + v1][wx + u1], z = px[2][wy + v1][wx + u1];
                    m0 += x; m1 += y; m2 += z;
                    // products of floats are exact in double: fma == add of the product
                    s00 = __builtin_fma(x, x, s00); s01 = __builtin_fma(x, y, s01); s02 = __builtin_fma(x, z, s02);
                    s11 = __builtin_fma(y, y, s11); s12 = __builtin_fma(y, z, s12); s22 = __builtin_fma(z, z, s22);
                }
            m0 *= kInv9; m1 *= kInv9; m2 *= kInv9;
            const double reg = (double)eps * kInv9;
            const double c11 = __builtin_fma(-m0, m0, s00 * kInv9) + reg;
            const double c12 = __builtin_fma(-m0, m1, s01 * kInv9);
            const double c13 = __builtin_fma(-m0, m2, s02 * kInv9);
            const double c22 = __builtin_fma(-m1, m1, s11 * kInv9) + reg;
            const double c23 = __builtin_fma(-m1, m2, s12 * kInv9);
            const double c33 = __builtin_fma(-m2, m2, s22 * kInv9) + reg;
            const double d11 = __builtin_fma(c22, c33, -(c23 * c23));
            const double d12 = __builtin_fma(c13, c23, -(c12 * c33));
            const double d13 = __builtin_fma(c12, c23, -(c13 * c22));
            const double d22 = __builtin_fma(c11, c33, -(c13 * c13));
            const double d23 = __builtin_fma(c12, c13, -(c11 * c23));
            const double d33 = __builtin_fma(c11, c22, -(c12 * c12));
            const double det = __builtin_fma(c11, d11, __builtin_fma(c12, d12, c13 * d13));
            const double q = -1.0 / (9.0 * det);
            ws[0][e] = d11 * q; ws[1][e] = d12 * q; ws[2][e] = d13 * q;
            ws[3][e] = d22 * q; ws[4][e] = d23 * q; ws[5][e] = d33 * q;
            const float h0 = (float)m0, h1 = (float)m1, h2 = (float)m2;
            wm[0][e] = h0; wm[1][e] = h1; wm[2][e] = h2;
            wm[3][e] = (float)(m0 - (double)h0);
            wm[4][e] = (float)(m1 - (double)h1);
            wm[5][e] = (float)(m2 - (double)h2);
        }
    }
    __syncthreads();

    const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= W || gy >= H) return;

    float acc[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) acc[k] = 0.f;
    const float xi0 = px[0][ty + 2][tx + 2], xi1 = px[1][ty + 2][tx + 2], xi2 = px[2][ty + 2][tx + 2];

    if constexpr (MODE == GM_ASSEMBLY_REFERENCE) {
#pragma unroll
        for (int wy2 = 0; wy2 < 3; ++wy2)
#pragma unroll
            for (int wx2 = 0; wx2 < 3; ++wx2) {
                const int e = (ty + wy2) * kWX + (tx + wx2);
                if (!wvalid[e]) continue;
                // Matting.cu:146-153
                const float mx = (float)ws[7][e], my = (float)ws[8][e], mz = (float)ws[9][e];
                const float rx = xi0 - mx, ry = xi1 - my, rz = xi2 - mz;
                const double det = ws[6][e];
                const float wx = (float)((ws[0][e] * (double)rx + ws[1][e] * (double)ry + ws[2][e] * (double)rz) / det);
                const float wy = (float)((ws[1][e] * (double)rx + ws[3][e] * (double)ry + ws[4][e] * (double)rz) / det);
                const float wz = (float)((ws[2][e] * (double)rx + ws[4][e] * (double)ry + ws[5][e] * (double)rz) / det);
#pragma unroll
                for (int vy = 0; vy < 3; ++vy)
#pragma unroll
                    for (int vx = 0; vx < 3; ++vx) {
                        const int lx = tx + wx2 + vx, ly = ty + wy2 + vy;  // pixel j in the halo
                        const float qx = px[0][ly][lx] - mx, qy = px[1][ly][lx] - my, qz = px[2][ly][lx] - mz;
                        // Matting.cu:177
                        float s = 1.0f + wx * qx;
                        s = s + wy * qy;
                        s = s + wz * qz;
                        acc[12 + (wx2 - 1 + vx - 1) + 5 * (wy2 - 1 + vy - 1)] += -s / 9.0f;
                    }
            }
    } else {
        // per window n (raster order; centre offset (n % 3 - 1, n / 3 - 1)):
        // v = S' (x_i - mu), c = -1/9 - v.ml.  An invalid window gets
        // v = c = 0: its terms are exactly +0 and leave the sums unchanged.
        float vc[9], v0[9], v1[9], v2[9], h0[9], h1[9], h2[9];
        const double xd0 = xi0, xd1 = xi1, xd2 = xi2;
#pragma unroll
        for (int n = 0; n < 9; ++n) {
            const int e = (ty + n / 3) * kWX + (tx + n % 3);
            vc[n] = v0[n] = v1[n] = v2[n] = h0[n] = h1[n] = h2[n] = 0.f;
            if (!wvalid[e]) continue;
            // x_i - mu in double, mu as its float head and tail
            const double r0 = (xd0 - (double)wm[0][e]) - (double)wm[3][e];
            const double r1 = (xd1 - (double)wm[1][e]) - (double)wm[4][e];
            const double r2 = (xd2 - (double)wm[2][e]) - (double)wm[5][e];
            const double a00 = ws[0][e], a01 = ws[1][e], a02 = ws[2][e];
            const double a11 = ws[3][e], a12 = ws[4][e], a22 = ws[5][e];
            v0[n] = (float)__builtin_fma(a02, r2, __builtin_fma(a01, r1, a00 * r0));
            v1[n] = (float)__builtin_fma(a12, r2, __builtin_fma(a11, r1, a01 * r0));
            v2[n] = (float)__builtin_fma(a22, r2, __builtin_fma(a12, r1, a02 * r0));
            vc[n] = __builtin_fmaf(-v2[n], wm[5][e],
                                   __builtin_fmaf(-v1[n], wm[4][e], __builtin_fmaf(-v0[n], wm[3][e], -kInv9f)));
            h0[n] = wm[0][e]; h1[n] = wm[1][e]; h2[n] = wm[2][e];
        }
        // member j = i + (dx, dy) lies in window n iff |dx - cx_n| <= 1 and
        // |dy - cy_n| <= 1; L[i, j] sums its windows in raster order and is
        // stored as soon as it is complete (the stores then overlap the rest
        // of the arithmetic, and the entry's register is free).  The diagonal
        // is set from the row sum below.
#pragma unroll
        for (int dy = -2; dy <= 2; ++dy)
#pragma unroll
            for (int dx = -2; dx <= 2; ++dx) {
                if (dx == 0 && dy == 0) continue;
                const int lx = tx + 2 + dx, ly = ty + 2 + dy;
                const float p0 = px[0][ly][lx], p1 = px[1][ly][lx], p2 = px[2][ly][lx];
#pragma unroll
                for (int n = 0; n < 9; ++n) {
                    const int cx = n % 3 - 1, cy = n / 3 - 1;
                    if (dx - cx < -1 || dx - cx > 1 || dy - cy < -1 || dy - cy > 1) continue;
                    const float t = __builtin_fmaf(
                        v2[n], p2 - h2[n], __builtin_fmaf(v1[n], p1 - h1[n], __builtin_fmaf(v0[n], p0 - h0[n], vc[n])));
                    acc[12 + dx + 5 * dy] += t;
                }
                A[(size_t)(12 + dx + 5 * dy) * apitch + (size_t)gy * (size_t)W + (size_t)gx] = acc[12 + dx + 5 * dy];
            }
    }

    const size_t i = (size_t)gy * (size_t)W + (size_t)gx;
    float l = 0.f, tg = 0.f;
    if (scr) {
        const unsigned char sv = scr[(size_t)gy * pitch + (size_t)gx];
        if (sv == 255) { l = gamma; tg = 1.f; }
        else if (sv == 0) { l = gamma; tg = 0.f; }
    } else if (lam) {
        l = lam[i];
        tg = tgt ? tgt[i] : 0.f;
    }
    if constexpr (MODE == GM_ASSEMBLY_REFERENCE) {
        // Matting.cu:191-218: L_ii -= rowsum (double), then the data term
        double rs = 0;
#pragma unroll
        for (int k = 0; k < 25; ++k) rs += (double)acc[k];
        acc[12] = (float)((double)acc[12] - rs);
    } else {
        double off = 0;
#pragma unroll
        for (int k = 0; k < 25; ++k)
            if (k != 12) off += (double)acc[k];
        acc[12] = (float)(-off);
    }
    float bi = 0.f;
    if (l != 0.f) {
        acc[12] += l;
        bi = l * tg;
    }
    if constexpr (MODE == GM_ASSEMBLY_REFERENCE) {
#pragma unroll
        for (int k = 0; k < 25; ++k) A[(size_t)k * apitch + i] = acc[k];
    } else {
        A[(size_t)12 * apitch + i] = acc[12];  // (the off-diagonal entries are stored)
    }
    if (bvec) bvec[i] = bi;
    if (dinv) dinv[i] = acc[12] != 0.f ? 1.f / acc[12] : 0.f;
}


// ---------------------------------------------------------------------------
// Assembly adjoint (differentiation path; no reference counterpart).
//
// For the system (L(I) + Lambda) alpha = Lambda t with adjoint mu = A^-1 dLoss/dalpha,
// dLoss/dL_ij = -mu_i alpha_j; with L_ii = -sum_{j!=i} L_ij + lam_i the
// effective weight of an off-diagonal entry is mu_i (alpha_i - alpha_j).  Per
// interior window (centred colours X, S = (X^T X/9 + eps/9 I)^-1, Y = X S,
// H_st = (mu_s - mu_t)(alpha_s - alpha_t), K = H Y):
//   dLoss/dX_p = -(1/9) [K_p - (1/9) Y_p (X^T K) - (1/9) 1^T K]
// summed over the windows containing pixel p (gather, window raster order).
// Everything in FP64 (exact double covariance, as GM_ASSEMBLY_ACCURATE).
// Also dLoss/dlam_i = mu_i (t_i - alpha_i) and dLoss/dt_i = mu_i lam_i.
constexpr int kBX = 32, kBY = 8;
constexpr int kBWX = kBX + 2, kBWY = kBY + 2, kBNW = kBWX * kBWY;
constexpr int kBPX = kBX + 4, kBPY = kBY + 4;

__global__ __launch_bounds__(kBlock) void k_levin_backward(const float* __restrict__ im, int nch, size_t pitch,
                                                           const float* __restrict__ mu,
                                                           const float* __restrict__ alpha,
                                                           const float* __restrict__ lam,
                                                           const float* __restrict__ tgt, float eps, int W, int H,
                                                           float* __restrict__ gim, float* __restrict__ glam,
                                                           float* __restrict__ gtgt) {
    __shared__ float px[5][kBPY][kBPX];  // r, g, b, mu, alpha
    __shared__ double wst[21][kBNW];     // S(6) m(3) XtK(9) 1tK(3)
    __shared__ unsigned char wok[kBNW];
    const int x0 = blockIdx.x * kBX, y0 = blockIdx.y * kBY;
    for (int e = threadIdx.x; e < kBPX * kBPY; e += kBlock) {
        const int ly = e / kBPX, lx = e % kBPX;
        const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
        float c0 = 0.f, c1 = 0.f, c2 = 0.f, m = 0.f, a = 0.f;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const float* p = im + ((size_t)gy * pitch + (size_t)gx) * (size_t)nch;
            c0 = p[0]; c1 = p[1]; c2 = p[2];
            const size_t i = (size_t)gy * W + gx;
            m = mu[i];
            a = alpha[i];
        }
        px[0][ly][lx] = c0; px[1][ly][lx] = c1; px[2][ly][lx] = c2; px[3][ly][lx] = m; px[4][ly][lx] = a;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kBNW; e += kBlock) {
        const int wy = e / kBWX, wx = e % kBWX;
        const int cx = x0 - 1 + wx, cy = y0 - 1 + wy;
        const bool ok = cx >= 1 && cx <= W - 2 && cy >= 1 && cy <= H - 2;
        wok[e] = ok;
        if (!ok) continue;
        double X[9][3], m0 = 0, m1 = 0, m2 = 0;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int ly = wy + t / 3, lx = wx + t % 3;
            X[t][0] = px[0][ly][lx]; X[t][1] = px[1][ly][lx]; X[t][2] = px[2][ly][lx];
            m0 += X[t][0]; m1 += X[t][1]; m2 += X[t][2];
        }
        m0 /= 9.0; m1 /= 9.0; m2 /= 9.0;
        double c11 = 0, c12 = 0, c13 = 0, c22 = 0, c23 = 0, c33 = 0;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            X[t][0] -= m0; X[t][1] -= m1; X[t][2] -= m2;
            c11 += X[t][0] * X[t][0]; c12 += X[t][0] * X[t][1]; c13 += X[t][0] * X[t][2];
            c22 += X[t][1] * X[t][1]; c23 += X[t][1] * X[t][2]; c33 += X[t][2] * X[t][2];
        }
        const double reg = (double)eps / 9.0;
        c11 = c11 / 9.0 + reg; c12 /= 9.0; c13 /= 9.0; c22 = c22 / 9.0 + reg; c23 /= 9.0; c33 = c33 / 9.0 + reg;
        const double d11 = c22 * c33 - c23 * c23, d12 = c13 * c23 - c12 * c33, d13 = c12 * c23 - c13 * c22;
        const double d22 = c11 * c33 - c13 * c13, d23 = c12 * c13 - c11 * c23, d33 = c11 * c22 - c12 * c12;
        const double det = c11 * d11 + c12 * d12 + c13 * d13;
        const double S[6] = {d11 / det, d12 / det, d13 / det, d22 / det, d23 / det, d33 / det};
        double Y[9][3];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            Y[t][0] = S[0] * X[t][0] + S[1] * X[t][1] + S[2] * X[t][2];
            Y[t][1] = S[1] * X[t][0] + S[3] * X[t][1] + S[4] * X[t][2];
            Y[t][2] = S[2] * X[t][0] + S[4] * X[t][1] + S[5] * X[t][2];
        }
        double XK[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, kk[3] = {0, 0, 0};
#pragma unroll
        for (int s = 0; s < 9; ++s) {
            const double ms = px[3][wy + s / 3][wx + s % 3], as = px[4][wy + s / 3][wx + s % 3];
            double K0 = 0, K1 = 0, K2 = 0;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const double h = (ms - (double)px[3][wy + t / 3][wx + t % 3]) * (as - (double)px[4][wy + t / 3][wx + t % 3]);
                K0 += h * Y[t][0]; K1 += h * Y[t][1]; K2 += h * Y[t][2];
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                XK[3 * r + 0] += X[s][r] * K0;
                XK[3 * r + 1] += X[s][r] * K1;
                XK[3 * r + 2] += X[s][r] * K2;
            }
            kk[0] += K0; kk[1] += K1; kk[2] += K2;
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) wst[q][e] = S[q];
        wst[6][e] = m0; wst[7][e] = m1; wst[8][e] = m2;
#pragma unroll
        for (int q = 0; q < 9; ++q) wst[9 + q][e] = XK[q];
        wst[18][e] = kk[0]; wst[19][e] = kk[1]; wst[20][e] = kk[2];
    }
    __syncthreads();
    const int tx = threadIdx.x % kBX, ty = threadIdx.x / kBX;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= W || gy >= H) return;
    double g0 = 0, g1 = 0, g2 = 0;
    const double mp = px[3][ty + 2][tx + 2], ap = px[4][ty + 2][tx + 2];
#pragma unroll
    for (int wy2 = 0; wy2 < 3; ++wy2)
#pragma unroll
        for (int wx2 = 0; wx2 < 3; ++wx2) {
            const int e = (ty + wy2) * kBWX + (tx + wx2);
            if (!wok[e]) continue;
            double S[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) S[q] = wst[q][e];
            const double m0 = wst[6][e], m1 = wst[7][e], m2 = wst[8][e];
            double K0 = 0, K1 = 0, K2 = 0;
#pragma unroll
            for (int vy = 0; vy < 3; ++vy)
#pragma unroll
                for (int vx = 0; vx < 3; ++vx) {
                    const int ly = ty + wy2 + vy, lx = tx + wx2 + vx;
                    const double x = (double)px[0][ly][lx] - m0, y = (double)px[1][ly][lx] - m1,
                                 z = (double)px[2][ly][lx] - m2;
                    const double h = (mp - (double)px[3][ly][lx]) * (ap - (double)px[4][ly][lx]);
                    K0 += h * (S[0] * x + S[1] * y + S[2] * z);
                    K1 += h * (S[1] * x + S[3] * y + S[4] * z);
                    K2 += h * (S[2] * x + S[4] * y + S[5] * z);
                }
            const double x = (double)px[0][ty + 2][tx + 2] - m0, y = (double)px[1][ty + 2][tx + 2] - m1,
                         z = (double)px[2][ty + 2][tx + 2] - m2;
            const double Y0 = S[0] * x + S[1] * y + S[2] * z;
            const double Y1 = S[1] * x + S[3] * y + S[4] * z;
            const double Y2 = S[2] * x + S[4] * y + S[5] * z;
            const double yk0 = Y0 * wst[9][e] + Y1 * wst[12][e] + Y2 * wst[15][e];
            const double yk1 = Y0 * wst[10][e] + Y1 * wst[13][e] + Y2 * wst[16][e];
            const double yk2 = Y0 * wst[11][e] + Y1 * wst[14][e] + Y2 * wst[17][e];
            g0 -= (K0 - yk0 / 9.0 - wst[18][e] / 9.0) / 9.0;
            g1 -= (K1 - yk1 / 9.0 - wst[19][e] / 9.0) / 9.0;
            g2 -= (K2 - yk2 / 9.0 - wst[20][e] / 9.0) / 9.0;
        }
    const size_t i = (size_t)gy * W + gx;
    gim[3 * i + 0] = (float)g0;
    gim[3 * i + 1] = (float)g1;
    gim[3 * i + 2] = (float)g2;
    if (glam) glam[i] = (float)(mp * ((tgt ? (double)tgt[i] : 0.0) - ap));
    if (gtgt) gtgt[i] = (float)(mp * (lam ? (double)lam[i] : 0.0));
}

// Assembly into a raw 25-band array (band k at A + k*apitch, rows W*H in raster
// order).  b and dinv may be NULL.
gm_status levin_assemble_raw(float* A, size_t apitch, float* b, float* dinv, const float* im, int nch, size_t pitch,
                             const float* lam, const float* tgt, const unsigned char* scr, float gamma, float eps,
                             int mode, int W, int H, gm_stream_t s) {
    if (!A || !im) return fail(GM_ERR_INVALID, "levin: null pointer");
    if (W < 1 || H < 1) return fail(GM_ERR_INVALID, "levin: image %dx%d", W, H);
    if (nch != 3 && nch != 4) return fail(GM_ERR_INVALID, "levin: nch must be 3 or 4");
    if (pitch < (size_t)W) return fail(GM_ERR_INVALID, "levin: pitch %zu < W %d", pitch, W);
    if (apitch < (size_t)W * H) return fail(GM_ERR_INVALID, "levin: apitch < W*H");
    if (mode != GM_ASSEMBLY_REFERENCE && mode != GM_ASSEMBLY_ACCURATE) return fail(GM_ERR_INVALID, "levin: mode");
    if (!(eps >= 0.f)) return fail(GM_ERR_INVALID, "levin: eps must be >= 0");
    dim3 grid((W + kTX - 1) / kTX, (H + kTY - 1) / kTY);
    if (grid.y > 65535u) return fail(GM_ERR_UNSUPPORTED, "levin: image too tall");
    if (mode == GM_ASSEMBLY_REFERENCE)
        hipLaunchKernelGGL(k_levin<GM_ASSEMBLY_REFERENCE>, grid, dim3(kBlock), 0, (hipStream_t)s, im, nch, pitch, lam,
                           tgt, scr, gamma, eps, W, H, A, apitch, b, dinv);
    else
        hipLaunchKernelGGL(k_levin<GM_ASSEMBLY_ACCURATE>, grid, dim3(kBlock), 0, (hipStream_t)s, im, nch, pitch, lam,
                           tgt, scr, gamma, eps, W, H, A, apitch, b, dinv);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static gm_status levin_launch(BandedMatrix* L, float* b, float* dinv, const float* im, int nch, size_t pitch,
                              const float* lam, const float* tgt, const unsigned char* scr, float gamma, float eps,
                              int mode, int W, int H, gm_stream_t s) {
    if (!L || !b || !im) return fail(GM_ERR_INVALID, "levin: null pointer");
    if (W < 1 || H < 1) return fail(GM_ERR_INVALID, "levin: image %dx%d", W, H);
    if (nch != 3 && nch != 4) return fail(GM_ERR_INVALID, "levin: nch must be 3 or 4");
    if (pitch < (size_t)W) return fail(GM_ERR_INVALID, "levin: pitch %zu < W %d", pitch, W);
    if ((long long)W * H != (long long)L->rows) return fail(GM_ERR_INVALID, "levin: L->rows != W*H");
    if (mode != GM_ASSEMBLY_REFERENCE && mode != GM_ASSEMBLY_ACCURATE) return fail(GM_ERR_INVALID, "levin: mode");
    if (!(eps >= 0.f)) return fail(GM_ERR_INVALID, "levin: eps must be >= 0");
    BandPlan p;
    gm_status st = get_plan(L, &p);
    if (st != GM_OK) return st;
    if (L->nbands != 25) return fail(GM_ERR_INVALID, "levin: L needs 25 bands");
    // the band order must be the 5x5 window order; for W < 5 the offsets alias
    // but the layout is still defined by main.cu:161-190
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx)
            if (p.args.off[5 * (dy + 2) + (dx + 2)] != dy * W + dx)
                return fail(GM_ERR_INVALID, "levin: L bands are not the 5x5 window offsets for W=%d", W);
    dim3 grid((W + kTX - 1) / kTX, (H + kTY - 1) / kTY);
    if (grid.y > 65535u) return fail(GM_ERR_UNSUPPORTED, "levin: image too tall");
    if (mode == GM_ASSEMBLY_REFERENCE)
        hipLaunchKernelGGL(k_levin<GM_ASSEMBLY_REFERENCE>, grid, dim3(kBlock), 0, (hipStream_t)s, im, nch, pitch, lam,
                           tgt, scr, gamma, eps, W, H, L->a, L->apitch, b, dinv);
    else
        hipLaunchKernelGGL(k_levin<GM_ASSEMBLY_ACCURATE>, grid, dim3(kBlock), 0, (hipStream_t)s, im, nch, pitch, lam,
                           tgt, scr, gamma, eps, W, H, L->a, L->apitch, b, dinv);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

}  // namespace gm

using namespace gm;

extern "C" {

gm_status gm_levin_laplacian(BandedMatrix* L, float* b, float* dinv, const float* im, int nch, size_t im_pitch,
                             const float* lam, const float* tgt, float eps, int mode, int W, int H, gm_stream_t s) {
    if (lam && !tgt) return fail(GM_ERR_INVALID, "levin: lam without tgt");
    return levin_launch(L, b, dinv, im, nch, im_pitch, lam, tgt, nullptr, 0.f, eps, mode, W, H, s);
}

gm_status gm_levin_laplacian_u8(BandedMatrix* L, float* b, float* dinv, const float* im, int nch, size_t im_pitch,
                                const unsigned char* scribs, float eps, float gamma, int mode, int W, int H,
                                gm_stream_t s) {
    if (!scribs) return fail(GM_ERR_INVALID, "levin: null scribbles");
    return levin_launch(L, b, dinv, im, nch, im_pitch, nullptr, nullptr, scribs, gamma, eps, mode, W, H, s);
}

gm_status gm_levin_backward(float* grad_im, float* grad_lam, float* grad_tgt, const float* im, int nch,
                             size_t im_pitch, const float* mu, const float* alpha, const float* lam, const float* tgt,
                             float eps, int W, int H, gm_stream_t s) {
    if (!grad_im || !im || !mu || !alpha) return fail(GM_ERR_INVALID, "levin_backward: null pointer");
    if (W < 1 || H < 1 || (nch != 3 && nch != 4) || im_pitch < (size_t)W)
        return fail(GM_ERR_INVALID, "levin_backward: bad image shape");
    dim3 grid((W + kBX - 1) / kBX, (H + kBY - 1) / kBY);
    if (grid.y > 65535u) return fail(GM_ERR_UNSUPPORTED, "levin_backward: image too tall");
    hipLaunchKernelGGL(k_levin_backward, grid, dim3(kBlock), 0, (hipStream_t)s, im, nch, im_pitch, mu, alpha, lam, tgt,
                       eps, W, H, grad_im, grad_lam, grad_tgt);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

gm_status gm_host_levin_laplacian(BandedMatrix* L, float* b, float eps, const float* im4,
                                  const unsigned char* scribs, int W, int H, size_t im_pitch) {
    if (!L || !L->a || !L->bands || !b || !im4 || !scribs) return fail(GM_ERR_INVALID, "hostLevinLaplacian: null");
    if (L->nbands != 25 || (long long)L->rows != (long long)W * H || (size_t)L->rows > L->apitch)
        return fail(GM_ERR_INVALID, "hostLevinLaplacian: L must have 25 bands and W*H rows");
    if (im_pitch < (size_t)W) return fail(GM_ERR_INVALID, "hostLevinLaplacian: pitch < W");
    BandedMatrix d{};
    float *db = nullptr, *dim = nullptr;
    unsigned char* dsc = nullptr;
    gm_status st = gm_bm_device_alloc(&d, L->rows, 25, L->bands);
    if (st != GM_OK) return st;
    const size_t npx = im_pitch * (size_t)(H > 0 ? H : 1);
    hipError_t e = hipMalloc(&db, sizeof(float) * (size_t)(L->rows > 0 ? L->rows : 1));
    if (e == hipSuccess) e = hipMalloc(&dim, sizeof(float) * 4 * npx);
    if (e == hipSuccess) e = hipMalloc(&dsc, npx);
    if (e == hipSuccess) e = hipMemcpy(dim, im4, sizeof(float) * 4 * npx, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dsc, scribs, npx, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        st = hip_fail(e, "hostLevinLaplacian staging", __FILE__, __LINE__);
    } else {
        // gamma = 1 as the reference hard-codes (Matting.cu:48)
        st = levin_launch(&d, db, nullptr, dim, 4, im_pitch, nullptr, nullptr, dsc, 1.0f, eps, GM_ASSEMBLY_REFERENCE,
                          W, H, nullptr);
        if (st == GM_OK) st = gm_bm_copy_to_host(L, &d);
        if (st == GM_OK && L->rows > 0) {
            e = hipMemcpy(b, db, sizeof(float) * L->rows, hipMemcpyDeviceToHost);
            if (e != hipSuccess) st = hip_fail(e, "hipMemcpy(b)", __FILE__, __LINE__);
        }
    }
    (void)hipFree(db);
    (void)hipFree(dim);
    (void)hipFree(dsc);
    gm_bm_device_free(&d);
    return st;
}

}  // extern "C"
