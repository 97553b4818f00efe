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

#endif  // GPUMATTING_HPP
