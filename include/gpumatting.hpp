// gpumatting.hpp -- the reference's host API under its own names, as thin C++
// wrappers over the C ABI (gpumatting.h).
//
// Code written against rocketman768/gpumatting keeps its calls: the names,
// parameter orders and default arguments below are those of BandedMatrix.cu,
// Vector.cu, Solve.h, Matting.cu and main.cu (file:line at each).  The only
// syntax change is that kernel launches (`bmAx_k<nb><<<g,b>>>(...)`) become
// plain calls (`bmAx(...)`) with an optional stream.  Like the reference these
// return void; the status of the last call on the thread is gmLastStatus()
// and its message gm_last_error().
#ifndef GPUMATTING_HPP
#define GPUMATTING_HPP

#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>

#include "gpumatting.h"

namespace gpumatting_detail {
inline int& last_status() {
    static thread_local int s = 0;
    return s;
}
inline void record(gm_status s) { last_status() = (int)s; }
}  // namespace gpumatting_detail

inline int gmLastStatus() { return gpumatting_detail::last_status(); }

// BandedMatrix.cu:23-41
inline void bmCopyToDevice(BandedMatrix* dA, BandedMatrix const* hA) {
    gpumatting_detail::record(gm_bm_copy_to_device(dA, hA));
}
// BandedMatrix.cu:43-47
inline void bmDeviceFree(BandedMatrix* dA) { gpumatting_detail::record(gm_bm_device_free(dA)); }
// bmAx_k, BandedMatrix.cu:107-141,194-198
inline void bmAx(float* b, const BandedMatrix a, float const* x, gm_stream_t s = 0) {
    gpumatting_detail::record(gm_bm_ax(b, &a, x, s));
}
// bmAxpy_k<nb, add>, BandedMatrix.cu:152-189,203-207
inline void bmAxpy(float* b, const BandedMatrix a, float const* x, float const* y, bool add, gm_stream_t s = 0) {
    gpumatting_detail::record(gm_bm_axpy(b, &a, x, y, add ? 1 : 0, s));
}
// jacobi_k, BandedMatrix.cu:61-98 (one sweep)
inline void jacobiSweep(float* xx, float const* x, const BandedMatrix a, float const* b, float omega,
                        gm_stream_t s = 0) {
    gpumatting_detail::record(gm_bm_jacobi_sweep(xx, x, &a, b, omega, s));
}

// Vector.cu:23-30
inline void vecDeviceMalloc(float** dx, int length, int leftPadding = 0, int rightPadding = 0) {
    gpumatting_detail::record(gm_vec_device_malloc(dx, length, leftPadding, rightPadding));
}
// Vector.cu:37-41
inline void vecCopyToDevice(float** dx, float const* hx, int length, int leftPadding = 0, int rightPadding = 0) {
    gpumatting_detail::record(gm_vec_copy_to_device(dx, hx, length, leftPadding, rightPadding));
}
// Vector.cu:43-46
inline void vecDeviceFree(float* dx, int leftPadding = 0) {
    gpumatting_detail::record(gm_vec_device_free(dx, leftPadding));
}
// vecAdd_k / vecSub_k, Vector.cu:142-150
inline void vecAdd(float* result, float const* x, float const* y, int len, gm_stream_t s = 0) {
    gpumatting_detail::record(gm_vec_add(result, x, y, len, s));
}
inline void vecSub(float* result, float const* x, float const* y, int len, gm_stream_t s = 0) {
    gpumatting_detail::record(gm_vec_sub(result, x, y, len, s));
}
// vecScaleConst_k, Vector.cu:175
inline void vecScaleConst(float* result, float const* x, float s, int len, gm_stream_t st = 0) {
    gpumatting_detail::record(gm_vec_scale_const(result, x, s, len, st));
}
// vecScale_k, Vector.cu:180 (device scalar)
inline void vecScale(float* result, float const* x, float* s, int len, gm_stream_t st = 0) {
    gpumatting_detail::record(gm_vec_scale(result, x, s, len, st));
}
// innerProd_k, Vector.cu:262 (device result)
inline void innerProd(float* result, float const* x, float const* y, int len, gm_stream_t s = 0) {
    gpumatting_detail::record(gm_inner_prod(result, x, y, len, s));
}

// Solve.h:17
inline void gaussSeidel_host(float* x, BandedMatrix const& A, float const* b, size_t iterations = 1) {
    gpumatting_detail::record(gm_gauss_seidel_host(x, &A, b, iterations));
}
// Solve.h:31
inline void jacobi_host(float* x, BandedMatrix const& A, float const* b, size_t iterations, int xpad,
                        float omega = 2.f / 3.f) {
    gpumatting_detail::record(gm_jacobi_host(x, &A, b, iterations, xpad, omega));
}

// main.cu:46
inline void gradSolve(float* alpha, BandedMatrix L, float* b, int iterations, int pad) {
    gpumatting_detail::record(gm_grad_solve(alpha, &L, b, iterations, pad, 0));
}
// main.cu:57
inline void cgSolve(float* alpha, BandedMatrix L, float* b, int pad, int iterations, int restartInterval) {
    gpumatting_detail::record(gm_cg_solve(alpha, &L, b, pad, iterations, restartInterval, 0));
}
// main.cu:77-84
inline void jacobi(float* x, const BandedMatrix a, float const* b, float omega, int pad, int iterations) {
    gpumatting_detail::record(gm_jacobi(x, &a, b, omega, pad, iterations, 0));
}

// Matting.cu:39-46.  `lambda` is the covariance regulariser eps (the
// reference's name); im is an RGBA float4 image (any 16-byte float4 type).
template <class Float4>
inline void hostLevinLaplacian(BandedMatrix L, float* b, float lambda, Float4* im, unsigned char* scribs, int imW,
                               int imH, size_t imPitch) {
    static_assert(sizeof(Float4) == 4 * sizeof(float), "im must be float4 pixels");
    gpumatting_detail::record(gm_host_levin_laplacian(&L, b, lambda, reinterpret_cast<const float*>(im), scribs, imW,
                                                      imH, imPitch));
}

// ---- sparse matrices and superpixels (CompressedMatrix.cu, CoordMatrix.cu,
// SLIC.h, ppm.h) -------------------------------------------------------------
// The reference's CSR/COO routines are device functions over device-side
// structs; here the structs live on the host and hold device arrays.
// CompressedMatrix.cu:20-31
inline void csmInit(CompressedMatrix* m, int rows, int cols, size_t length) {
    gpumatting_detail::record(gm_csm_init(m, rows, cols, length));
}
// csmAx, CompressedMatrix.cu:118-165: b = A x
inline void csmAx(float* b, CompressedMatrix const* a, float const* x, gm_stream_t s = 0) {
    gpumatting_detail::record(gm_csm_ax(b, a, x, s));
}
// csmAxpy, CompressedMatrix.cu:45-85: b = A x + y
inline void csmAxpy(float* b, CompressedMatrix const* a, float const* x, float const* y, gm_stream_t s = 0) {
    gpumatting_detail::record(gm_csm_axpy(b, a, x, y, 1, s));
}
// CoordMatrix.cu:20-29
inline void cooInit(CoordMatrix* m, int rows, int cols, size_t length) {
    gpumatting_detail::record(gm_coo_init(m, rows, cols, length));
}
// CoordMatrix.cu:31-40
inline void cooRelease(CoordMatrix* m) { gpumatting_detail::record(gm_coo_release(m)); }
// CoordMatrix.cu:42-48 (host entries; the reference's `int* k` holds the
// values, which are float)
inline void cooAppend(CoordMatrix* m, int* i, int* j, float* k, size_t length) {
    gpumatting_detail::record(gm_coo_append(m, i, j, k, length));
}
// cooAxpy, CoordMatrix.cu:50-87: b += A x + y
inline void cooAxpy(float* b, CoordMatrix const* a, float const* x, float const* y, gm_stream_t s = 0) {
    gpumatting_detail::record(gm_coo_axpy(b, a, x, y, s));
}
// ppm.h:151-165: 8-bit RGB to 4 bytes per pixel A, R, G, B (A = 0xff); malloc'd
inline unsigned char* rgbToArgb(unsigned char* rgb, int w, int h) {
    unsigned char* ret = (unsigned char*)malloc((size_t)4 * w * h);
    if (!ret) return NULL;
    for (size_t p = 0; p < (size_t)w * h; ++p) {
        ret[4 * p] = 0xff;
        ret[4 * p + 1] = rgb[3 * p];
        ret[4 * p + 2] = rgb[3 * p + 1];
        ret[4 * p + 3] = rgb[3 * p + 2];
    }
    return ret;
}
// SLIC.h:43: host pixels laid out as rgbToArgb makes them, host labels out
// (width * height ints); returns the number of superpixels (0 on failure,
// status in gmLastStatus()).  nseg is the requested count K, and
// spatialConsistency the compactness m.
inline int slicSegmentation(int* labels, const unsigned int* in, int width, int height, int nseg,
                            double spatialConsistency = 10.0) {
    const int n = width * height;
    float* h = (float*)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    if (!h) {
        gpumatting_detail::record(GM_ERR_OOM);
        return 0;
    }
    const unsigned char* px = reinterpret_cast<const unsigned char*>(in);
    for (int p = 0; p < n; ++p)
        for (int c = 0; c < 3; ++c) h[3 * p + c] = px[4 * p + 1 + c] / 255.0f;
    float *dim = NULL, *dlab = NULL;
    int nl = 0;
    gm_status st = gm_vec_copy_to_device(&dim, h, 3 * n, 0, 0);
    if (st == GM_OK) st = gm_vec_device_malloc(&dlab, n, 0, 0);
    if (st == GM_OK)
        st = gm_slic(reinterpret_cast<int*>(dlab), &nl, dim, 3, (size_t)width, width, height, nseg,
                     (float)spatialConsistency, 10, 0);
    if (st == GM_OK) st = gm_vec_copy_to_host(reinterpret_cast<float*>(labels), dlab, n);
    if (dim) gm_vec_device_free(dim, 0);
    if (dlab) gm_vec_device_free(dlab, 0);
    free(h);
    gpumatting_detail::record(st);
    return st == GM_OK ? nl : 0;
}

// ---- driver I/O (ppm.h, main.cu) ------------------------------------------
// Like the reference, readers return malloc'd buffers (release with free())
// or NULL on failure (message in gm_last_error()), writers return 0 / -1.

struct gm_float4 {
    float x, y, z, w;
};

// ppm.h:32-129
inline unsigned char* pgmread(const char* filename, int* w, int* h) {
    unsigned char* d = nullptr;
    const gm_status s = gm_pgm_read(filename, &d, w, h, nullptr);
    gpumatting_detail::record(s);
    if (s != GM_OK) *w = *h = 0;
    return s == GM_OK ? d : nullptr;
}
// ppm.h:134-149: samples / 255
inline float* pgmread_float(const char* filename, int* w, int* h) {
    unsigned char* c = pgmread(filename, w, h);
    if (!c) return nullptr;
    const size_t n = (size_t)(*w) * (*h);
    float* f = (float*)malloc(n * sizeof(float));
    if (f)
        for (size_t i = 0; i < n; ++i) f[i] = (float)c[i] / 255.0f;
    free(c);
    return f;
}
// ppm.h:178-278
inline unsigned char* ppmread(const char* filename, int* w, int* h, int* maxval) {
    unsigned char* d = nullptr;
    const gm_status s = gm_ppm_read(filename, &d, w, h, maxval);
    gpumatting_detail::record(s);
    if (s != GM_OK) *w = *h = 0;
    return s == GM_OK ? d : nullptr;
}
// ppm.h:285-302: samples / maxval, 3 per pixel
inline float* ppmread_float(const char* filename, int* w, int* h) {
    int maxval = 0;
    unsigned char* c = ppmread(filename, w, h, &maxval);
    if (!c) return nullptr;
    const size_t n = (size_t)(*w) * (*h) * 3;
    float* f = (float*)malloc(n * sizeof(float));
    if (f)
        for (size_t i = 0; i < n; ++i) f[i] = (float)c[i] / (float)maxval;
    free(c);
    return f;
}
// ppm.h:309-337: RGBA pixels, samples / maxval, alpha 1; the 8-bit samples
// are handed back in *outcdata when it is not NULL
template <class Float4 = gm_float4>
inline Float4* ppmread_float4(unsigned char** outcdata, const char* filename, int* w, int* h) {
    static_assert(sizeof(Float4) == 4 * sizeof(float), "float4 pixels");
    int maxval = 0;
    unsigned char* c = ppmread(filename, w, h, &maxval);
    if (!c) return nullptr;
    const size_t n = (size_t)(*w) * (*h);
    float* f = (float*)malloc(n * 4 * sizeof(float));
    if (f)
        for (size_t i = 0; i < n; ++i) {
            f[4 * i + 0] = (float)c[3 * i + 0] / (float)maxval;
            f[4 * i + 1] = (float)c[3 * i + 1] / (float)maxval;
            f[4 * i + 2] = (float)c[3 * i + 2] / (float)maxval;
            f[4 * i + 3] = 1.0f;
        }
    if (outcdata) *outcdata = c;
    else free(c);
    return reinterpret_cast<Float4*>(f);
}
// ppm.h:345-418
inline int pgmwrite(const char* filename, int w, int h, const unsigned char* data, const char* comment_string,
                    int binsave) {
    const gm_status s = gm_pgm_write(filename, data, w, h, comment_string, binsave);
    gpumatting_detail::record(s);
    return s == GM_OK ? 0 : -1;
}
// ppm.h:430-451
inline int pgmwrite_float(const char* filename, int w, int h, const float* data, const char* comment_string,
                          int binsave) {
    const gm_status s = gm_pgm_write_float(filename, data, w, h, comment_string, binsave);
    gpumatting_detail::record(s);
    return s == GM_OK ? 0 : -1;
}
// main.cu:558-577: prints "Ground truth MSE: <mse>"
inline void computeError(const float* alpha, const float* gtAlpha, int imW, int imH) {
    double mse = 0.0;
    const gm_status s = gm_clamped_mse(alpha, gtAlpha, imW * imH, &mse);
    gpumatting_detail::record(s);
    if (s == GM_OK) printf("Ground truth MSE: %.3e\n", mse);
}
// main.cu:334-355: the row count and band offsets to stderr, then the host
// matrix to A.bin in the working directory
inline void dump(BandedMatrix const& bm) {
    fprintf(stderr, "%d\n", bm.rows);
    for (int i = 0; i < bm.nbands; ++i) fprintf(stderr, "%d,", bm.bands[i]);
    printf("\n");
    gpumatting_detail::record(gm_bm_dump(&bm, "A.bin"));
}

#endif  // GPUMATTING_HPP
