/*
 * oracle.c -- CPU restatement of gpumatting's closed-form matting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load the
 * library built from it.  The product (gpumatting_amd/, libgpumatting.so)
 * never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math, so the
 * float/double mix of the reference is reproduced operation by operation).
 *
 * Parity pins (what checks each function independently of its own formula):
 *   orc_levin_scatter_f32 : dense brute-force Levin (orc_levin_dense_f64, a
 *                           different inverse: Gauss-Jordan) to 1e-5 rel; row
 *                           sums; symmetry; constant-image closed form;
 *                           3x3-image single-window case.
 *   orc_levin_gather_f32  : bit-identical to orc_levin_scatter_f32.
 *   orc_levin_f64         : dense brute force to 1e-10.
 *   orc_bm_ax_f32         : BandedMatrixTest.cu 3-band known answer; dense
 *                           mat-vec.
 *   orc_dot_f32           : VectorTest.cu known answer; math.fsum.
 *   orc_cg_ref_f32 / orc_gd_ref_f32 / orc_pcg_* : numpy.linalg.solve on small
 *                           SPD systems; monotone A-norm error for CG.
 *   orc_gs_f32 / orc_jacobi_f32 : reference Solve.cpp compiled into
 *                           oracle/_ref (even iteration counts) and a dense
 *                           textbook sweep.
 *   orc_levin_backward_f64: central finite differences of orc_levin_f64 +
 *                           orc_pcg_f64.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NB 25
#define CB 12 /* centre band */

/* ------------------------------------------------------------------------ */
/* Band layout: main.cu:161-190.  band k = 5(dy+2)+(dx+2) <-> offset dy*W+dx. */
void orc_bands25(int W, int* bands)
{
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx)
            bands[5 * (dy + 2) + (dx + 2)] = dy * W + dx;
}

/* Vector padding: main.cu:228-229. */
int orc_pad(int W) { return ((2 * W + 2) / 32) * 32 + 32; }

/* ------------------------------------------------------------------------ */
/* Per-window statistics exactly as Matting.cu:95-137 computes them.
 * mu in float (float4 accumulation, :96,:103,:114); covariance sums in double
 * of float products (:104-109); eps/|w| added in float then promoted
 * (:115-120); determinant and adjugate in double (:130-137). */
typedef struct {
    float mx, my, mz;
    double d11, d12, d13, d22, d23, d33, det;
} orc_win_t;

static void orc_window_stats(orc_win_t* w, const float* im, int nch, size_t pitch,
                             int u, int v, float eps)
{
    float mx = 0.f, my = 0.f, mz = 0.f;
    double c11 = 0, c12 = 0, c13 = 0, c22 = 0, c23 = 0, c33 = 0;
    const int n = 9;
    for (int v1 = v - 1; v1 <= v + 1; ++v1)
        for (int u1 = u - 1; u1 <= u + 1; ++u1) {
            const float* px = im + ((size_t)v1 * pitch + (size_t)u1) * (size_t)nch;
            float x = px[0], y = px[1], z = px[2];
            mx += x; my += y; mz += z;
            float t;
            t = x * x; c11 += t;
            t = x * y; c12 += t;
            t = x * z; c13 += t;
            t = y * y; c22 += t;
            t = y * z; c23 += t;
            t = z * z; c33 += t;
        }
    mx /= (float)n; my /= (float)n; mz /= (float)n;
    float reg = eps / (float)n;
    float t;
    t = mx * mx; c11 = c11 / n - (double)t + (double)reg;
    t = mx * my; c12 = c12 / n - (double)t;
    t = mx * mz; c13 = c13 / n - (double)t;
    t = my * my; c22 = c22 / n - (double)t + (double)reg;
    t = my * mz; c23 = c23 / n - (double)t;
    t = mz * mz; c33 = c33 / n - (double)t + (double)reg;
    w->det = c11 * c22 * c33 + c12 * c23 * c13 + c13 * c12 * c23
           - c13 * c22 * c13 - c12 * c12 * c33 - c11 * c23 * c23;
    w->d11 = c33 * c22 - c23 * c23;
    w->d12 = c13 * c23 - c33 * c12;
    w->d13 = c23 * c12 - c22 * c13;
    w->d22 = c33 * c11 - c13 * c13;
    w->d23 = c12 * c13 - c23 * c11;
    w->d33 = c22 * c11 - c12 * c12;
    w->mx = mx; w->my = my; w->mz = mz;
}

/* Whitened pixel, Matting.cu:146-153: double products, rounded to float. */
static void orc_whiten(const orc_win_t* w, const float* px, float* wx, float* wy, float* wz,
                       float* rx, float* ry, float* rz)
{
    float x = px[0] - w->mx, y = px[1] - w->my, z = px[2] - w->mz;
    *rx = x; *ry = y; *rz = z;
    *wx = (float)((w->d11 * x + w->d12 * y + w->d13 * z) / w->det);
    *wy = (float)((w->d12 * x + w->d22 * y + w->d23 * z) / w->det);
    *wz = (float)((w->d13 * x + w->d23 * y + w->d33 * z) / w->det);
}

/* negSim, Matting.cu:171-177, all float. */
static inline float orc_negsim(float wx, float wy, float wz, const orc_win_t* w, const float* pj)
{
    float x = pj[0] - w->mx, y = pj[1] - w->my, z = pj[2] - w->mz;
    float s = 1.0f + wx * x;
    s = s + wy * y;
    s = s + wz * z;
    return -s / 9.0f;
}

/* Diagonal fix + data term, Matting.cu:191-218, generalised from the u8 scribble
 * rule (255 -> +gamma, b=gamma; 0 -> +gamma, b=0; else b=0) to per-pixel weights
 * lam[i] and targets tgt[i]: L_ii += lam_i (when lam_i != 0), b_i = lam_i*tgt_i.
 * lam == NULL means no data term. */
static void orc_finish_row(float* a, size_t apitch, float* b, size_t i,
                           const float* lam, const float* tgt, size_t li)
{
    double rowsum = 0;
    for (int k = 0; k < NB; ++k) rowsum += a[i + (size_t)k * apitch];
    a[i + CB * apitch] = (float)((double)a[i + CB * apitch] - rowsum);
    float l = lam ? lam[li] : 0.f;
    if (l != 0.f) {
        a[i + CB * apitch] += l;
        b[i] = l * tgt[li];
    } else {
        b[i] = 0.f;
    }
}

/*
 * Reference-semantics FP32 Levin Laplacian, scatter form: Matting.cu:39-219
 * restated.  Interior windows (u in [1,W-2], v in [1,H-2]) visited in raster
 * order, pixels i and j of each window in raster order, L[i, 12+du+5dv] +=
 * negSim.  a must have 25*apitch floats; it is zeroed here (the reference
 * requires a pre-zeroed L, main.cu:192-193).  im: nch (3 or 4) interleaved
 * floats per pixel, pitch in pixels; lam/tgt: W*H dense or NULL.
 */
void orc_levin_scatter_f32(float* a, size_t apitch, float* b, const float* im, int nch,
                           size_t pitch, const float* lam, const float* tgt, float eps,
                           int W, int H)
{
    for (int k = 0; k < NB; ++k) memset(a + (size_t)k * apitch, 0, sizeof(float) * (size_t)W * H);
    for (int v = 1; v < H - 1; ++v)
        for (int u = 1; u < W - 1; ++u) {
            orc_win_t w;
            orc_window_stats(&w, im, nch, pitch, u, v, eps);
            for (int v1 = v - 1; v1 <= v + 1; ++v1)
                for (int u1 = u - 1; u1 <= u + 1; ++u1) {
                    size_t i = (size_t)u1 + (size_t)v1 * W;
                    float wx, wy, wz, rx, ry, rz;
                    orc_whiten(&w, im + ((size_t)v1 * pitch + u1) * nch, &wx, &wy, &wz, &rx, &ry, &rz);
                    for (int v2 = v - 1; v2 <= v + 1; ++v2)
                        for (int u2 = u - 1; u2 <= u + 1; ++u2) {
                            int band = CB + (u2 - u1) + 5 * (v2 - v1);
                            float ns = orc_negsim(wx, wy, wz, &w, im + ((size_t)v2 * pitch + u2) * nch);
                            a[i + (size_t)band * apitch] += ns;
                        }
                }
        }
    for (int v = 0; v < H; ++v)
        for (int u = 0; u < W; ++u)
            orc_finish_row(a, apitch, b, (size_t)u + (size_t)v * W, lam, tgt, (size_t)u + (size_t)v * W);
}

/* u8-scribble front end with the reference's own rule, Matting.cu:205-216
 * (scribbles indexed with the image pitch, as the reference does at :205). */
void orc_levin_scatter_u8(float* a, size_t apitch, float* b, const float* im, int nch,
                          size_t pitch, const unsigned char* scribs, float eps, float gamma,
                          int W, int H)
{
    size_t n = (size_t)W * H;
    float* lam = (float*)calloc(n, sizeof(float));
    float* tgt = (float*)calloc(n, sizeof(float));
    for (int v = 0; v < H; ++v)
        for (int u = 0; u < W; ++u) {
            unsigned char s = scribs[(size_t)u + (size_t)v * pitch];
            size_t i = (size_t)u + (size_t)v * W;
            if (s == 255) { lam[i] = gamma; tgt[i] = 1.f; }
            else if (s == 0) { lam[i] = gamma; tgt[i] = 0.f; }
        }
    orc_levin_scatter_f32(a, apitch, b, im, nch, pitch, lam, tgt, eps, W, H);
    free(lam);
    free(tgt);
}

/*
 * Gather form of the same FP32 assembly: row i collects, for each interior
 * window containing it (window centres in raster order), the 9 contributions
 * of that window.  Each entry L[i,band] therefore receives the same float
 * additions in the same order as the scatter form -> bit-identical output.
 * Row-parallel with OpenMP (this is also the CPU baseline's assembly).
 * Writes every band entry (no pre-zeroing needed) and, if dinv != NULL,
 * dinv[i] = 1/L_ii (0 where L_ii == 0).
 */
void orc_levin_gather_f32(float* a, size_t apitch, float* b, float* dinv, const float* im,
                          int nch, size_t pitch, const float* lam, const float* tgt,
                          float eps, int W, int H)
{
#pragma omp parallel
    {
    /* window statistics of the three window rows touching pixel row py */
    orc_win_t* wst = (orc_win_t*)malloc(sizeof(orc_win_t) * 3 * (size_t)(W > 0 ? W : 1));
#pragma omp for schedule(static)
    for (int py = 0; py < H; ++py) {
        for (int cy = py - 1; cy <= py + 1; ++cy)
            for (int cx = 1; cx < W - 1; ++cx)
                if (cy >= 1 && cy <= H - 2)
                    orc_window_stats(&wst[(size_t)(cy - py + 1) * W + cx], im, nch, pitch, cx, cy, eps);
        for (int px = 0; px < W; ++px) {
            float acc[NB];
            for (int k = 0; k < NB; ++k) acc[k] = 0.f;
            const float* pi = im + ((size_t)py * pitch + px) * nch;
            for (int cy = py - 1; cy <= py + 1; ++cy)
                for (int cx = px - 1; cx <= px + 1; ++cx) {
                    if (cy < 1 || cy > H - 2 || cx < 1 || cx > W - 2) continue;
                    const orc_win_t* w = &wst[(size_t)(cy - py + 1) * W + cx];
                    float wx, wy, wz, rx, ry, rz;
                    orc_whiten(w, pi, &wx, &wy, &wz, &rx, &ry, &rz);
                    for (int vy = cy - 1; vy <= cy + 1; ++vy)
                        for (int vx = cx - 1; vx <= cx + 1; ++vx) {
                            int band = CB + (vx - px) + 5 * (vy - py);
                            acc[band] += orc_negsim(wx, wy, wz, w, im + ((size_t)vy * pitch + vx) * nch);
                        }
                }
            size_t i = (size_t)px + (size_t)py * W;
            for (int k = 0; k < NB; ++k) a[i + (size_t)k * apitch] = acc[k];
            orc_finish_row(a, apitch, b, i, lam, tgt, i);
            if (dinv) {
                float d = a[i + CB * apitch];
                dinv[i] = d != 0.f ? 1.f / d : 0.f;
            }
        }
    }
    free(wst);
    }
}

/* ------------------------------------------------------------------------ */
/* "Accurate" FP32 assembly (not the reference's rounding; same definition).
 * The reference forms the window covariance from float means and float
 * products (Matting.cu:96-120), which loses ~4 digits when the colour variance
 * is ~1e-4 and eps is 1e-7.  Here, per window: exact double sums of the
 * colours and of their products (float products are exact in double), the
 * 1/9 scalings as multiplications by the double 1/9, the covariance as
 * fma(-mu_a, mu_b, s_ab/9) (+ eps/9 on the diagonal), FMA-contracted adjugate
 * and determinant, and S' = -inverse/9 = adjugate * (-1 / (9 det)); the mean
 * mu as mh + ml (float head and tail).  Per pixel i and window w (raster
 * order): r = (x_i - mh) - ml (double), v = float(S' r) (double FMA chains)
 * and c = -1/9 - v.ml (float FMAs); every other member j = i + (dx, dy) adds
 * c + v.(x_j - mh) (float FMAs, x_j - mh in float) to L[i, j], windows in
 * raster order.  The diagonal is -sum of the off-diagonal entries in double,
 * then + lam_i.  Every fused operation is an explicit fma()/fmaf() (the file
 * is built without contraction), so this is the GPU kernel's arithmetic
 * operation for operation. */
typedef struct {
    double m[3];
    double S[6]; /* S'11 S'12 S'13 S'22 S'23 S'33 (= -inverse / 9) */
    float mh[3], ml[3]; /* mean = mh + ml: mh = float(mean), ml = float(mean - mh) */
} orc_wacc_t;

static const double ORC_INV9 = 1.0 / 9.0;

static void orc_window_stats_acc(orc_wacc_t* w, const float* im, int nch, size_t pitch,
                                 int u, int v, float eps)
{
    double m0 = 0, m1 = 0, m2 = 0, s00 = 0, s01 = 0, s02 = 0, s11 = 0, s12 = 0, s22 = 0;
    for (int v1 = v - 1; v1 <= v + 1; ++v1)
        for (int u1 = u - 1; u1 <= u + 1; ++u1) {
            const float* p = im + ((size_t)v1 * pitch + u1) * nch;
            double x = p[0], y = p[1], z = p[2];
            m0 += x; m1 += y; m2 += z;
            s00 = fma(x, x, s00); s01 = fma(x, y, s01); s02 = fma(x, z, s02);
            s11 = fma(y, y, s11); s12 = fma(y, z, s12); s22 = fma(z, z, s22);
        }
    m0 *= ORC_INV9; m1 *= ORC_INV9; m2 *= ORC_INV9;
    double reg = (double)eps * ORC_INV9;
    double c11 = fma(-m0, m0, s00 * ORC_INV9) + reg;
    double c12 = fma(-m0, m1, s01 * ORC_INV9);
    double c13 = fma(-m0, m2, s02 * ORC_INV9);
    double c22 = fma(-m1, m1, s11 * ORC_INV9) + reg;
    double c23 = fma(-m1, m2, s12 * ORC_INV9);
    double c33 = fma(-m2, m2, s22 * ORC_INV9) + reg;
    double d11 = fma(c22, c33, -(c23 * c23));
    double d12 = fma(c13, c23, -(c12 * c33));
    double d13 = fma(c12, c23, -(c13 * c22));
    double d22 = fma(c11, c33, -(c13 * c13));
    double d23 = fma(c12, c13, -(c11 * c23));
    double d33 = fma(c11, c22, -(c12 * c12));
    double det = fma(c11, d11, fma(c12, d12, c13 * d13));
    double q = -1.0 / (9.0 * det);
    w->S[0] = d11 * q; w->S[1] = d12 * q; w->S[2] = d13 * q;
    w->S[3] = d22 * q; w->S[4] = d23 * q; w->S[5] = d33 * q;
    w->m[0] = m0; w->m[1] = m1; w->m[2] = m2;
    for (int c = 0; c < 3; ++c) {
        w->mh[c] = (float)w->m[c];
        w->ml[c] = (float)(w->m[c] - (double)w->mh[c]);
    }
}

void orc_levin_gather_acc_f32(float* a, size_t apitch, float* b, float* dinv, const float* im,
                              int nch, size_t pitch, const float* lam, const float* tgt,
                              float eps, int W, int H)
{
#pragma omp parallel
    {
    orc_wacc_t* wst = (orc_wacc_t*)malloc(sizeof(orc_wacc_t) * 3 * (size_t)(W > 0 ? W : 1));
#pragma omp for schedule(static)
    for (int py = 0; py < H; ++py) {
        for (int cy = py - 1; cy <= py + 1; ++cy)
            for (int cx = 1; cx < W - 1; ++cx)
                if (cy >= 1 && cy <= H - 2)
                    orc_window_stats_acc(&wst[(size_t)(cy - py + 1) * W + cx], im, nch, pitch, cx, cy, eps);
        for (int px = 0; px < W; ++px) {
            float acc[NB];
            for (int k = 0; k < NB; ++k) acc[k] = 0.f;
            const float* pi = im + ((size_t)py * pitch + px) * nch;
            for (int cy = py - 1; cy <= py + 1; ++cy)
                for (int cx = px - 1; cx <= px + 1; ++cx) {
                    if (cy < 1 || cy > H - 2 || cx < 1 || cx > W - 2) continue;
                    const orc_wacc_t* w = &wst[(size_t)(cy - py + 1) * W + cx];
                    const double* S = w->S;
                    /* x_i - mu in double, mu as its float head and tail */
                    double r0 = ((double)pi[0] - (double)w->mh[0]) - (double)w->ml[0];
                    double r1 = ((double)pi[1] - (double)w->mh[1]) - (double)w->ml[1];
                    double r2 = ((double)pi[2] - (double)w->mh[2]) - (double)w->ml[2];
                    double u0 = fma(S[2], r2, fma(S[1], r1, S[0] * r0));
                    double u1 = fma(S[4], r2, fma(S[3], r1, S[1] * r0));
                    double u2 = fma(S[5], r2, fma(S[4], r1, S[2] * r0));
                    float v0 = (float)u0, v1 = (float)u1, v2 = (float)u2;
                    float c = fmaf(-v2, w->ml[2], fmaf(-v1, w->ml[1], fmaf(-v0, w->ml[0], -1.0f / 9.0f)));
                    for (int vy = cy - 1; vy <= cy + 1; ++vy)
                        for (int vx = cx - 1; vx <= cx + 1; ++vx) {
                            if (vx == px && vy == py) continue; /* the diagonal comes from the row sum */
                            const float* pj = im + ((size_t)vy * pitch + vx) * nch;
                            float q0 = pj[0] - w->mh[0], q1 = pj[1] - w->mh[1], q2 = pj[2] - w->mh[2];
                            acc[CB + (vx - px) + 5 * (vy - py)] += fmaf(v2, q2, fmaf(v1, q1, fmaf(v0, q0, c)));
                        }
                }
            size_t i = (size_t)px + (size_t)py * W;
            double off = 0;
            for (int k = 0; k < NB; ++k)
                if (k != CB) off += (double)acc[k];
            acc[CB] = (float)(-off);
            float l = lam ? lam[i] : 0.f;
            if (l != 0.f) {
                acc[CB] += l;
                b[i] = l * tgt[i];
            } else {
                b[i] = 0.f;
            }
            for (int k = 0; k < NB; ++k) a[i + (size_t)k * apitch] = acc[k];
            if (dinv) dinv[i] = acc[CB] != 0.f ? 1.f / acc[CB] : 0.f;
        }
    }
    free(wst);
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 Levin Laplacian (the accuracy oracle): same definition as above
 * (Matting.cu:39-219 / getLevinLapRaster.m:31-62) with every operation in
 * double.  Gather form, OpenMP over rows. */
static void orc_window_stats64(double* mu, double* S, const float* im, int nch, size_t pitch,
                               int u, int v, double eps)
{
    double m[3] = {0, 0, 0}, c[6] = {0, 0, 0, 0, 0, 0};
    for (int v1 = v - 1; v1 <= v + 1; ++v1)
        for (int u1 = u - 1; u1 <= u + 1; ++u1) {
            const float* p = im + ((size_t)v1 * pitch + u1) * nch;
            double x = p[0], y = p[1], z = p[2];
            m[0] += x; m[1] += y; m[2] += z;
            c[0] += x * x; c[1] += x * y; c[2] += x * z;
            c[3] += y * y; c[4] += y * z; c[5] += z * z;
        }
    for (int q = 0; q < 3; ++q) m[q] /= 9.0;
    double c11 = c[0] / 9 - m[0] * m[0] + eps / 9, c12 = c[1] / 9 - m[0] * m[1];
    double c13 = c[2] / 9 - m[0] * m[2], c22 = c[3] / 9 - m[1] * m[1] + eps / 9;
    double c23 = c[4] / 9 - m[1] * m[2], c33 = c[5] / 9 - m[2] * m[2] + eps / 9;
    double det = c11 * (c22 * c33 - c23 * c23) - c12 * (c12 * c33 - c23 * c13)
               + c13 * (c12 * c23 - c22 * c13);
    S[0] = (c22 * c33 - c23 * c23) / det; /* S11 */
    S[1] = (c13 * c23 - c12 * c33) / det; /* S12 */
    S[2] = (c12 * c23 - c13 * c22) / det; /* S13 */
    S[3] = (c11 * c33 - c13 * c13) / det; /* S22 */
    S[4] = (c12 * c13 - c11 * c23) / det; /* S23 */
    S[5] = (c11 * c22 - c12 * c12) / det; /* S33 */
    mu[0] = m[0]; mu[1] = m[1]; mu[2] = m[2];
}

/* lam64/tgt64 may be NULL.  Produces (L + Lambda) in a, Lambda*tgt in b. */
void orc_levin_f64(double* a, size_t apitch, double* b, const float* im, int nch, size_t pitch,
                   const double* lam, const double* tgt, double eps, int W, int H)
{
#pragma omp parallel for schedule(static)
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            double acc[NB];
            for (int k = 0; k < NB; ++k) acc[k] = 0.0;
            const float* pi = im + ((size_t)py * pitch + px) * nch;
            for (int cy = py - 1; cy <= py + 1; ++cy)
                for (int cx = px - 1; cx <= px + 1; ++cx) {
                    if (cy < 1 || cy > H - 2 || cx < 1 || cx > W - 2) continue;
                    double mu[3], S[6];
                    orc_window_stats64(mu, S, im, nch, pitch, cx, cy, eps);
                    double x = pi[0] - mu[0], y = pi[1] - mu[1], z = pi[2] - mu[2];
                    double wx = S[0] * x + S[1] * y + S[2] * z;
                    double wy = S[1] * x + S[3] * y + S[4] * z;
                    double wz = S[2] * x + S[4] * y + S[5] * z;
                    for (int vy = cy - 1; vy <= cy + 1; ++vy)
                        for (int vx = cx - 1; vx <= cx + 1; ++vx) {
                            if (vx == px && vy == py) continue;
                            const float* pj = im + ((size_t)vy * pitch + vx) * nch;
                            double s = 1.0 + wx * (pj[0] - mu[0]) + wy * (pj[1] - mu[1])
                                     + wz * (pj[2] - mu[2]);
                            acc[CB + (vx - px) + 5 * (vy - py)] -= s / 9.0;
                        }
                }
            double diag = 0;
            for (int k = 0; k < NB; ++k)
                if (k != CB) diag -= acc[k];
            size_t i = (size_t)px + (size_t)py * W;
            double l = lam ? lam[i] : 0.0;
            acc[CB] = diag + l;
            for (int k = 0; k < NB; ++k) a[i + (size_t)k * apitch] = acc[k];
            b[i] = lam ? l * tgt[i] : 0.0;
        }
}

/* 3x3 inverse by Gauss-Jordan elimination with partial pivoting (deliberately a
 * different method from the adjugate used by the reference and the kernels). */
static void orc_inv3_gj(const double* M, double* Minv)
{
    double A[3][6];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 6; ++c) A[r][c] = c < 3 ? M[3 * r + c] : (c - 3 == r ? 1.0 : 0.0);
    for (int c = 0; c < 3; ++c) {
        int p = c;
        for (int r = c + 1; r < 3; ++r)
            if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
        if (p != c)
            for (int q = 0; q < 6; ++q) { double t = A[c][q]; A[c][q] = A[p][q]; A[p][q] = t; }
        double piv = A[c][c];
        for (int q = 0; q < 6; ++q) A[c][q] /= piv;
        for (int r = 0; r < 3; ++r)
            if (r != c) {
                double f = A[r][c];
                for (int q = 0; q < 6; ++q) A[r][q] -= f * A[c][q];
            }
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Minv[3 * r + c] = A[r][c + 3];
}

/* Dense brute-force Levin Laplacian, getLevinLapRaster.m:31-62 restated
 * (raster index i = u + v*W): A(win,win) += (1 + Xc inv(Cov + eps/9 I) Xc^T)/9,
 * L = diag(rowsum A) - A.  L is N x N row-major, N = W*H (small images only). */
void orc_levin_dense_f64(double* L, const float* im, int nch, size_t pitch, double eps, int W, int H)
{
    size_t N = (size_t)W * H;
    memset(L, 0, sizeof(double) * N * N);
    double* A = (double*)calloc(N * N, sizeof(double));
    for (int v = 1; v < H - 1; ++v)
        for (int u = 1; u < W - 1; ++u) {
            int idx[9];
            double X[9][3], mu[3] = {0, 0, 0};
            int q = 0;
            /* column-major window order as the .m file; order is irrelevant here */
            for (int du = -1; du <= 1; ++du)
                for (int dv = -1; dv <= 1; ++dv) {
                    const float* p = im + ((size_t)(v + dv) * pitch + (u + du)) * nch;
                    idx[q] = (u + du) + (v + dv) * W;
                    for (int c = 0; c < 3; ++c) { X[q][c] = p[c]; mu[c] += p[c]; }
                    ++q;
                }
            for (int c = 0; c < 3; ++c) mu[c] /= 9.0;
            double C[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double s = 0;
                    for (int t = 0; t < 9; ++t) s += X[t][r] * X[t][c];
                    C[3 * r + c] = s / 9.0 - mu[r] * mu[c] + (r == c ? eps / 9.0 : 0.0);
                }
            double Ci[9];
            orc_inv3_gj(C, Ci);
            for (int t = 0; t < 9; ++t)
                for (int c = 0; c < 3; ++c) X[t][c] -= mu[c];
            for (int s = 0; s < 9; ++s)
                for (int t = 0; t < 9; ++t) {
                    double val = 0;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) val += X[s][r] * Ci[3 * r + c] * X[t][c];
                    A[(size_t)idx[s] * N + idx[t]] += (1.0 + val) / 9.0;
                }
        }
    for (size_t i = 0; i < N; ++i) {
        double rs = 0;
        for (size_t j = 0; j < N; ++j) rs += A[i * N + j];
        for (size_t j = 0; j < N; ++j) L[i * N + j] = -A[i * N + j];
        L[i * N + i] += rs;
    }
    free(A);
}

/* The diagonal band: index nbands/2 (Solve.cpp:10) when its offset is 0, else
 * the first band with offset 0. */
static int orc_center_band(const int* bands, int nbands)
{
    if (nbands > 0 && bands[nbands / 2] == 0) return nbands / 2;
    for (int k = 0; k < nbands; ++k)
        if (bands[k] == 0) return k;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Banded SpMV, BandedMatrix.cu:148-182: b_i = sum_k a[i+k*apitch]*x[i+bands[k]],
 * float accumulation in band order.  x must be readable at i+bands[k]
 * (padded, BandedMatrix.cu:145-146). */
void orc_bm_ax_f32(float* y, const float* a, size_t apitch, const int* bands, int nbands,
                   int rows, const float* x)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rows; ++i) {
        float s = 0.f;
        for (int k = 0; k < nbands; ++k) s += a[(size_t)i + (size_t)k * apitch] * x[(ptrdiff_t)i + bands[k]];
        y[i] = s;
    }
}

/* b = A x + y (add=1) or A x - y (add=0), BandedMatrix.cu:193-230. */
void orc_bm_axpy_f32(float* out, const float* a, size_t apitch, const int* bands, int nbands,
                     int rows, const float* x, const float* y, int add)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rows; ++i) {
        float s = 0.f;
        for (int k = 0; k < nbands; ++k) s += a[(size_t)i + (size_t)k * apitch] * x[(ptrdiff_t)i + bands[k]];
        out[i] = add ? s + y[i] : s - y[i];
    }
}

/* Rows in blocks of 256, bands outer within a block: every row still sums its
 * products in band order from 0.0, so the result is that of the plain row
 * loop, and the inner loop vectorizes. */
__attribute__((optimize("tree-vectorize")))
void orc_bm_ax_f64(double* y, const double* a, size_t apitch, const int* bands, int nbands,
                   int rows, const double* x)
{
#pragma omp parallel for schedule(static)
    for (int i0 = 0; i0 < rows; i0 += 256) {
        const int n = rows - i0 < 256 ? rows - i0 : 256;
        double s[256];
        for (int t = 0; t < n; ++t) s[t] = 0.0;
        for (int k = 0; k < nbands; ++k) {
            const double* ak = a + (size_t)k * apitch + i0;
            const double* xk = x + (ptrdiff_t)i0 + bands[k];
            for (int t = 0; t < n; ++t) s[t] += ak[t] * xk[t];
        }
        for (int t = 0; t < n; ++t) y[i0 + t] = s[t];
    }
}

/* Inner product, Vector.cu:195-260 (known answer VectorTest.cu:36-65).  The
 * reference accumulates float products in float with a nondeterministic
 * atomicAdd order; here float products are summed in double in index order. */
double orc_dot_f32(const float* x, const float* y, int n)
{
    double s = 0;
    for (int i = 0; i < n; ++i) s += (double)(x[i] * y[i]);
    return s;
}

/* FP64 sums over fixed chunks (ORC_NCHUNK of them, partial sums added in chunk
 * order), so the result does not depend on the thread count. */
#define ORC_NCHUNK 256
static size_t orc_chunk_lo(size_t n, int c) { return n * (size_t)c / ORC_NCHUNK; }

static double orc_dot_f64(const double* x, const double* y, int n)
{
    double part[ORC_NCHUNK];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ORC_NCHUNK; ++c) {
        double s = 0;
        for (size_t i = orc_chunk_lo((size_t)n, c); i < orc_chunk_lo((size_t)n, c + 1); ++i) s += x[i] * y[i];
        part[c] = s;
    }
    double s = 0;
    for (int c = 0; c < ORC_NCHUNK; ++c) s += part[c];
    return s;
}

/* ------------------------------------------------------------------------ */
/* Conjugate gradient, main.cu:458-526: r = L alpha - b; p = -r; per iteration
 * Lp = L p, k = <r,r>/<p,Lp>, alpha += k p, r += k Lp, k = <r,r>_new/<r,r>_old,
 * p = k p - r; r is recomputed from alpha (restart) every restartInterval
 * iterations, starting with the first.  Fixed iteration count.  alpha, r and p
 * are padded by `pad` (main.cu:471-473); FP32 vectors, dot products summed in
 * double (see orc_dot_f32). */
void orc_cg_ref_f32(float* alpha, const float* a, size_t apitch, const int* bands, int nbands,
                    int rows, const float* b, int pad, int iterations, int restart)
{
    size_t n = (size_t)rows, len = n + 2 * (size_t)pad;
    float* rb = (float*)calloc(len, sizeof(float));
    float* pb = (float*)calloc(len, sizeof(float));
    float* Lp = (float*)calloc(n, sizeof(float));
    float *r = rb + pad, *p = pb + pad;
    int inner = 0;
    while (iterations-- > 0) {
        if (inner == 0) {
            orc_bm_axpy_f32(r, a, apitch, bands, nbands, rows, alpha, b, 0);
            for (size_t i = 0; i < n; ++i) p[i] = -1.0f * r[i];
            inner = restart - 1;
        } else {
            --inner;
        }
        orc_bm_ax_f32(Lp, a, apitch, bands, nbands, rows, p);
        float rTr = (float)orc_dot_f32(r, r, rows);
        float pLp = (float)orc_dot_f32(p, Lp, rows);
        float k = rTr / pLp;
        for (size_t i = 0; i < n; ++i) alpha[i] = alpha[i] + k * p[i];
        for (size_t i = 0; i < n; ++i) r[i] = r[i] + k * Lp[i];
        k = (float)orc_dot_f32(r, r, rows);
        k = k / rTr;
        for (size_t i = 0; i < n; ++i) p[i] = k * p[i] - r[i];
    }
    free(rb); free(pb); free(Lp);
}

/* Steepest descent with exact line search, main.cu:406-456:
 * d = L alpha - b; k = (<d,b> - <Ld,alpha>)/<Ld,d>; alpha += k d. */
void orc_gd_ref_f32(float* alpha, const float* a, size_t apitch, const int* bands, int nbands,
                    int rows, const float* b, int pad, int iterations)
{
    size_t n = (size_t)rows;
    float* db = (float*)calloc(n + 2 * (size_t)pad, sizeof(float));
    float* e = (float*)calloc(n, sizeof(float));
    float* d = db + pad;
    while (iterations-- > 0) {
        orc_bm_axpy_f32(d, a, apitch, bands, nbands, rows, alpha, b, 0);
        float k = (float)orc_dot_f32(d, b, rows);
        orc_bm_ax_f32(e, a, apitch, bands, nbands, rows, d);
        float t = (float)orc_dot_f32(e, alpha, rows);
        k = k - t;
        t = (float)orc_dot_f32(e, d, rows);
        k = k / t;
        for (size_t i = 0; i < n; ++i) alpha[i] = alpha[i] + k * d[i];
    }
    free(db); free(e);
}

/* Jacobi-preconditioned CG (not in the reference; added on the same matrix,
 * diagonal from the band with offset 0).  Same sign convention as
 * orc_cg_ref_f32 (r = A x - b).  Stops when ||r||/||b|| <= tol (recursive
 * residual) or after maxit iterations; restart > 0 recomputes r from x every
 * `restart` iterations.  Returns iterations done; *relres = final recursive
 * ||r||/||b||.  precond = 0 gives plain CG with the same stopping rule. */
int orc_pcg_f32(float* x, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                const float* b, int pad, double tol, int maxit, int restart, int precond,
                double* relres)
{
    size_t n = (size_t)rows, len = n + 2 * (size_t)pad;
    int cb = orc_center_band(bands, nbands);
    float* rb = (float*)calloc(len, sizeof(float));
    float* pb = (float*)calloc(len, sizeof(float));
    float* Ap = (float*)calloc(n, sizeof(float));
    float* dinv = (float*)calloc(n, sizeof(float));
    float *r = rb + pad, *p = pb + pad;
    for (size_t i = 0; i < n; ++i) {
        float d = a[i + (size_t)cb * apitch];
        dinv[i] = (precond && d != 0.f) ? 1.f / d : 1.f;
    }
    double bn = sqrt(orc_dot_f32(b, b, rows));
    if (bn == 0) bn = 1;
    double rz = 0, rr = 0;
    int it = 0, since = 0;
    int fresh = 1;
    for (;;) {
        if (fresh) {
            orc_bm_axpy_f32(r, a, apitch, bands, nbands, rows, x, b, 0);
            rz = 0; rr = 0;
            for (size_t i = 0; i < n; ++i) {
                float z = dinv[i] * r[i];
                rz += (double)(r[i] * z);
                rr += (double)(r[i] * r[i]);
                p[i] = -z;
            }
            fresh = 0;
            since = 0;
        }
        if (sqrt(rr) / bn <= tol || it >= maxit) break;
        orc_bm_ax_f32(Ap, a, apitch, bands, nbands, rows, p);
        double pAp = orc_dot_f32(p, Ap, rows);
        if (!(pAp > 0)) break;
        float al = (float)(rz / pAp);
        double rzn = 0;
        rr = 0;
        for (size_t i = 0; i < n; ++i) {
            x[i] = x[i] + al * p[i];
            r[i] = r[i] + al * Ap[i];
            float z = dinv[i] * r[i];
            rzn += (double)(r[i] * z);
            rr += (double)(r[i] * r[i]);
        }
        float be = (float)(rzn / rz);
        rz = rzn;
        for (size_t i = 0; i < n; ++i) p[i] = -(dinv[i] * r[i]) + be * p[i];
        ++it;
        ++since;
        if (restart > 0 && since >= restart) fresh = 1;
    }
    if (relres) *relres = sqrt(rr) / bn;
    free(rb); free(pb); free(Ap); free(dinv);
    return it;
}

/* FP64 Jacobi-PCG on an FP64 banded matrix (the accuracy oracle's solver). */
int orc_pcg_f64(double* x, const double* a, size_t apitch, const int* bands, int nbands, int rows,
                const double* b, int pad, double tol, int maxit, double* relres)
{
    size_t n = (size_t)rows, len = n + 2 * (size_t)pad;
    int cb = orc_center_band(bands, nbands);
    double* rb = (double*)calloc(len, sizeof(double));
    double* pb = (double*)calloc(len, sizeof(double));
    double* Ap = (double*)calloc(n, sizeof(double));
    double* dinv = (double*)calloc(n, sizeof(double));
    double *r = rb + pad, *p = pb + pad;
    for (size_t i = 0; i < n; ++i) {
        double d = a[i + (size_t)cb * apitch];
        dinv[i] = d != 0 ? 1.0 / d : 1.0;
    }
    double bn = sqrt(orc_dot_f64(b, b, rows));
    if (bn == 0) bn = 1;
    orc_bm_ax_f64(r, a, apitch, bands, nbands, rows, x);
    double rz = 0, rr = 0;
    for (size_t i = 0; i < n; ++i) {
        r[i] -= b[i];
        double z = dinv[i] * r[i];
        rz += r[i] * z;
        rr += r[i] * r[i];
        p[i] = -z;
    }
    int it = 0;
    double prz[ORC_NCHUNK], prr[ORC_NCHUNK];
    while (sqrt(rr) / bn > tol && it < maxit) {
        orc_bm_ax_f64(Ap, a, apitch, bands, nbands, rows, p);
        double pAp = orc_dot_f64(p, Ap, rows);
        if (!(pAp > 0)) break;
        double al = rz / pAp, rzn = 0;
        rr = 0;
#pragma omp parallel for schedule(static)
        for (int c = 0; c < ORC_NCHUNK; ++c) {
            double sz = 0, sr = 0;
            for (size_t i = orc_chunk_lo(n, c); i < orc_chunk_lo(n, c + 1); ++i) {
                x[i] += al * p[i];
                r[i] += al * Ap[i];
                sz += r[i] * dinv[i] * r[i];
                sr += r[i] * r[i];
            }
            prz[c] = sz;
            prr[c] = sr;
        }
        for (int c = 0; c < ORC_NCHUNK; ++c) {
            rzn += prz[c];
            rr += prr[c];
        }
        double be = rzn / rz;
        rz = rzn;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < rows; ++i) p[i] = -dinv[i] * r[i] + be * p[i];
        ++it;
    }
    if (relres) *relres = sqrt(rr) / bn;
    free(rb); free(pb); free(Ap); free(dinv);
    return it;
}

/* True relative residual ||b - A x|| / ||b|| with A in float, x float, summed in
 * double (the FP64 refinement residual definition). */
double orc_true_relres_f32(const float* a, size_t apitch, const int* bands, int nbands, int rows,
                           const float* b, const float* x)
{
    double num = 0, den = 0;
#pragma omp parallel for reduction(+ : num, den) schedule(static)
    for (int i = 0; i < rows; ++i) {
        double s = 0;
        for (int k = 0; k < nbands; ++k)
            s += (double)a[(size_t)i + (size_t)k * apitch] * (double)x[(ptrdiff_t)i + bands[k]];
        double d = (double)b[i] - s;
        num += d * d;
        den += (double)b[i] * (double)b[i];
    }
    return den > 0 ? sqrt(num / den) : sqrt(num);
}

/* Squared FP64 residual of one batched system, as the FP32 system a
 * BandedMatrix holds (Matting.cu:205-216 adds the constraint weight to the
 * diagonal in float): A = L except a_ii = fl(L_ii + lam_i), b_i = fl(lam_i *
 * tgt_i).  L has no data term; x is padded by the band reach.  out[0] =
 * ||A x - b||^2, out[1] = ||b||^2, both summed in FP64 over fixed chunks. */
void orc_resid_data_f64(double* out, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                        const float* lam, const float* tgt, const float* x)
{
    int cb = orc_center_band(bands, nbands);
    double pn[ORC_NCHUNK], pd[ORC_NCHUNK];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ORC_NCHUNK; ++c) {
        double sn = 0, sd = 0;
        for (size_t i = orc_chunk_lo((size_t)rows, c); i < orc_chunk_lo((size_t)rows, c + 1); ++i) {
            double s = 0;
            for (int k = 0; k < nbands; ++k) {
                float v = a[i + (size_t)k * apitch];
                if (k == cb) v = v + lam[i];
                s += (double)v * (double)x[(ptrdiff_t)i + bands[k]];
            }
            double bi = (double)(lam[i] * tgt[i]);
            sn += (s - bi) * (s - bi);
            sd += bi * bi;
        }
        pn[c] = sn;
        pd[c] = sd;
    }
    out[0] = out[1] = 0;
    for (int c = 0; c < ORC_NCHUNK; ++c) {
        out[0] += pn[c];
        out[1] += pd[c];
    }
}

/* ------------------------------------------------------------------------ */
/* Host Gauss-Seidel, Solve.cpp:3-33: in-place forward sweep skipping the band
 * with offset 0 in the sum (Solve.cpp:22-26), x_i = (b_i - sum)/a_ii. */
void orc_gs_f32(float* x, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                const float* b, size_t iterations)
{
    int mid = nbands / 2;
    for (; iterations > 0; --iterations)
        for (int i = 0; i < rows; ++i) {
            float t = 0.f;
            for (int k = 0; k < nbands; ++k)
                if (k != mid) t += a[(size_t)i + (size_t)k * apitch] * x[(ptrdiff_t)i + bands[k]];
            t = b[i] - t;
            x[i] = t / a[(size_t)i + (size_t)mid * apitch];
        }
}

/* Host damped Jacobi, Solve.cpp:35-85: xx_i = omega*(b_i - sum_{k!=mid})/a_ii
 * + (1-omega)*x_i.  Unlike the reference (which swaps buffers and, after an
 * odd number of iterations, leaves the result in its internal buffer and frees
 * the caller's, Solve.cpp:53-63), the result always ends up in x. */
void orc_jacobi_f32(float* x, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                    const float* b, size_t iterations, int xpad, float omega)
{
    int mid = nbands / 2;
    float om1 = 1.f - omega;
    float* xb = (float*)calloc((size_t)rows + 2 * (size_t)xpad, sizeof(float));
    float* xx = xb + xpad;
    for (; iterations > 0; --iterations) {
        for (int i = 0; i < rows; ++i) {
            float t = 0.f;
            for (int k = 0; k < nbands; ++k)
                if (k != mid) t += a[(size_t)i + (size_t)k * apitch] * x[(ptrdiff_t)i + bands[k]];
            t = b[i] - t;
            xx[i] = omega * t / a[(size_t)i + (size_t)mid * apitch] + om1 * x[i];
        }
        memcpy(x, xx, sizeof(float) * (size_t)rows);
    }
    free(xb);
}

/* Clamped ground-truth MSE, main.cu:558-577. */
double orc_mse_clamped(const float* alpha, const float* gt, int W, int H)
{
    double ssd = 0;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            float a = alpha[j + i * W], g = gt[j + i * W];
            if (a > 1.0f) ssd += (1.0f - g) * (1.0f - g);
            else if (a < 0.0f) ssd += g * g;
            else ssd += (a - g) * (a - g);
        }
    return ssd / ((double)W * H);
}

/* OpenMP team size for parallel regions the calling thread starts (a
 * caller running several oracle calls on its own threads sets 1). */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* CPU baseline (timed by bench.py, not a checker): the same Jacobi-PCG as
 * orc_pcg_f32 with every vector loop OpenMP-parallel (static schedule, one
 * fused pass per step), FP32 vectors, double reductions.  Iterates from x
 * until the recursive ||r||/||b|| <= tol (tol <= 0: exactly maxit iterations)
 * or maxit iterations; *iters gets the count.  Returns the final ||r||/||b||. */
double orc_baseline_pcg_f32(float* x, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                            const float* b, int pad, int maxit, double tol, int* iters)
{
    size_t n = (size_t)rows, len = n + 2 * (size_t)pad;
    int cb = orc_center_band(bands, nbands);
    float* rb = (float*)calloc(len, sizeof(float));
    float* pb = (float*)calloc(len, sizeof(float));
    float* Ap = (float*)calloc(n > 0 ? n : 1, sizeof(float));
    float *r = rb + pad, *p = pb + pad;
    const float* dg = a + (size_t)cb * apitch;
    double bb = 0, rz = 0, rr = 0;
#pragma omp parallel for reduction(+ : bb, rz, rr) schedule(static)
    for (int i = 0; i < rows; ++i) {
        float s = 0.f;
        for (int k = 0; k < nbands; ++k) s += a[(size_t)i + (size_t)k * apitch] * x[(ptrdiff_t)i + bands[k]];
        s -= b[i];
        r[i] = s;
        float z = dg[i] != 0.f ? s / dg[i] : s;
        p[i] = -z;
        rz += (double)(s * z);
        rr += (double)(s * s);
        bb += (double)b[i] * b[i];
    }
    const double tol2 = tol > 0 ? tol * tol * (bb > 0 ? bb : 1.0) : -1.0;
    int it = 0;
    while (it < maxit && !(rr <= tol2)) {
        double pAp = 0;
#pragma omp parallel for reduction(+ : pAp) schedule(static)
        for (int i = 0; i < rows; ++i) {
            float s = 0.f;
            for (int k = 0; k < nbands; ++k) s += a[(size_t)i + (size_t)k * apitch] * p[(ptrdiff_t)i + bands[k]];
            Ap[i] = s;
            pAp += (double)(p[i] * s);
        }
        if (!(pAp > 0)) break;
        float al = (float)(rz / pAp);
        double rzn = 0;
        rr = 0;
#pragma omp parallel for reduction(+ : rzn, rr) schedule(static)
        for (int i = 0; i < rows; ++i) {
            x[i] += al * p[i];
            float ri = r[i] + al * Ap[i];
            r[i] = ri;
            float z = dg[i] != 0.f ? ri / dg[i] : ri;
            rzn += (double)(ri * z);
            rr += (double)(ri * ri);
        }
        float be = (float)(rzn / rz);
        rz = rzn;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < rows; ++i) {
            float z = dg[i] != 0.f ? r[i] / dg[i] : r[i];
            p[i] = -z + be * p[i];
        }
        ++it;
    }
    if (iters) *iters = it;
    free(rb); free(pb); free(Ap);
    return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the assembly (FP64; no reference counterpart: the
 * differentiation path).  For a loss whose gradient with respect to the
 * stored off-diagonal entries is G_ij (the diagonal being -sum of the
 * off-diagonals + lam), and G_ij = -mu_i alpha_j for a linear solve
 * (L + Lambda) alpha = Lambda t with adjoint mu, the effective weight of entry
 * (i,j), i != j, is gbar_ij = mu_i (alpha_i - alpha_j).  Per interior window w
 * with centred colours X (9x3), S = (X^T X/9 + eps/9 I)^-1, Y = X S and
 * H_ij = (mu_i - mu_j)(alpha_i - alpha_j):
 *   dLoss/dX = Z - 1 (1^T Z)/9,  Z = -(1/9) [ H Y - (1/9) Y (X^T H Y) ]
 * and the rows of dLoss/dX are added to the window's 9 pixels.
 * grad (W*H*3 doubles, channel-interleaved) is overwritten. */
void orc_levin_backward_f64(double* grad, const float* im, int nch, size_t pitch, const double* mu,
                            const double* alpha, double eps, int W, int H)
{
    memset(grad, 0, sizeof(double) * (size_t)W * H * 3);
    for (int v = 1; v < H - 1; ++v)
        for (int u = 1; u < W - 1; ++u) {
            int idx[9];
            double X[9][3], m[3] = {0, 0, 0};
            int q = 0;
            for (int dv = -1; dv <= 1; ++dv)
                for (int du = -1; du <= 1; ++du) {
                    const float* p = im + ((size_t)(v + dv) * pitch + (u + du)) * nch;
                    idx[q] = (u + du) + (v + dv) * W;
                    for (int c = 0; c < 3; ++c) { X[q][c] = p[c]; m[c] += p[c]; }
                    ++q;
                }
            for (int c = 0; c < 3; ++c) m[c] /= 9.0;
            for (int t = 0; t < 9; ++t)
                for (int c = 0; c < 3; ++c) X[t][c] -= m[c];
            double C[9], S[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double s = 0;
                    for (int t = 0; t < 9; ++t) s += X[t][r] * X[t][c];
                    C[3 * r + c] = s / 9.0 + (r == c ? eps / 9.0 : 0.0);
                }
            orc_inv3_gj(C, S);
            double Y[9][3], K[9][3], Hm[9][9];
            for (int t = 0; t < 9; ++t)
                for (int c = 0; c < 3; ++c) {
                    double s = 0;
                    for (int r = 0; r < 3; ++r) s += X[t][r] * S[3 * r + c];
                    Y[t][c] = s;
                }
            for (int s = 0; s < 9; ++s)
                for (int t = 0; t < 9; ++t)
                    Hm[s][t] = (mu[idx[s]] - mu[idx[t]]) * (alpha[idx[s]] - alpha[idx[t]]);
            for (int s = 0; s < 9; ++s)
                for (int c = 0; c < 3; ++c) {
                    double acc = 0;
                    for (int t = 0; t < 9; ++t) acc += Hm[s][t] * Y[t][c];
                    K[s][c] = acc;
                }
            double XK[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double acc = 0;
                    for (int t = 0; t < 9; ++t) acc += X[t][r] * K[t][c];
                    XK[r][c] = acc;
                }
            double Z[9][3], colsum[3] = {0, 0, 0};
            for (int t = 0; t < 9; ++t)
                for (int c = 0; c < 3; ++c) {
                    double yk = 0;
                    for (int r = 0; r < 3; ++r) yk += Y[t][r] * XK[r][c];
                    Z[t][c] = -(K[t][c] - yk / 9.0) / 9.0;
                    colsum[c] += Z[t][c];
                }
            for (int t = 0; t < 9; ++t)
                for (int c = 0; c < 3; ++c) grad[(size_t)idx[t] * 3 + c] += Z[t][c] - colsum[c] / 9.0;
        }
}
