// gm_smooth.hip -- relaxation sweeps: damped Jacobi (jacobi main.cu:528-556,
// jacobi_host Solve.cpp:35-85) and Gauss-Seidel (gaussSeidel_host
// Solve.cpp:3-33).
//
// Gauss-Seidel is sequential by definition; the GPU sweep keeps the exact
// lexicographic semantics (every row reads the new values of earlier rows and
// the old values of later rows) by a hyperplane ordering:
//   * Levin 5x5 structure: pixel (u, v) runs at step t = u + 3v.  Its "new"
//     neighbours (u+2, v-1), (u-1, v) ... have t' <= t-1 and its "old"
//     neighbours (u-2, v+1), (u+1, v) ... have t' >= t+1, and two pixels of
//     one step are never stencil neighbours, so each step is a parallel batch.
//   * any other band set: rows are processed in chunks of c consecutive rows,
//     c = min |nonzero offset|, chunks in order (rows of a chunk never read
//     each other).
// The per-row arithmetic follows Solve.cpp exactly (float multiply then add in
// band order skipping the band with offset 0, then (b - sum) / a_ii, no FMA
// contraction), so the GPU sweep reproduces the host reference bit for bit.
#include <vector>

#include "gm_internal.h"

namespace gm {

// Solve.cpp:20-29 / 41-50 row update; `mid` is the index of the band with offset 0.
__device__ __forceinline__ float gs_row(const float* __restrict__ a, size_t pitch, const BandArgs& ba, int mid,
                                        int i, const float* x, const float* __restrict__ b) {
#pragma clang fp contract(off)
    float t = 0.f;
    for (int k = 0; k < ba.nb; ++k) {
        if (k == mid) continue;
        const float p = a[(size_t)i + (size_t)k * pitch] * x[(long long)i + ba.off[k]];
        t = t + p;
    }
    t = b[i] - t;
    return t / a[(size_t)i + (size_t)mid * pitch];
}

// one workgroup, Levin wavefront; x is updated in place
__global__ __launch_bounds__(1024) void k_gs_wavefront(float* x, const float* __restrict__ a, size_t pitch, BandArgs ba,
                                                       int mid, int W, int H, const float* __restrict__ b,
                                                       int iterations) {
    const int steps = (W - 1) + 3 * (H - 1) + 1;
    for (int it = 0; it < iterations; ++it) {
        for (int t = 0; t < steps; ++t) {
            int vlo = t - (W - 1) > 0 ? (t - (W - 1) + 2) / 3 : 0;
            int vhi = t / 3 < H - 1 ? t / 3 : H - 1;
            for (int v = vlo + (int)threadIdx.x; v <= vhi; v += blockDim.x) {
                const int u = t - 3 * v;
                const int i = u + v * W;
                const float nx = gs_row(a, pitch, ba, mid, i, x, b);
                x[i] = nx;
            }
            __syncthreads();
        }
    }
}

// one workgroup, chunks of c independent rows
__global__ __launch_bounds__(1024) void k_gs_chunks(float* x, const float* __restrict__ a, size_t pitch, BandArgs ba,
                                                    int mid, int rows, int c, const float* __restrict__ b,
                                                    int iterations) {
    for (int it = 0; it < iterations; ++it) {
        for (int i0 = 0; i0 < rows; i0 += c) {
            const int i1 = min(rows, i0 + c);
            for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
                const float nx = gs_row(a, pitch, ba, mid, i, x, b);
                x[i] = nx;
            }
            __syncthreads();
        }
    }
}

// damped Jacobi with Solve.cpp:41-50 arithmetic: xx = omega*(b - sum_{k!=mid})/a_ii + (1-omega)*x
__global__ __launch_bounds__(kBlock) void k_jacobi_host_form(float* __restrict__ xx, const float* __restrict__ x,
                                                             const float* __restrict__ a, size_t pitch, BandArgs ba,
                                                             int mid, int rows, const float* __restrict__ b,
                                                             float omega) {
#pragma clang fp contract(off)
    const float om1 = 1.f - omega;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < rows; i += gridDim.x * kBlock) {
        float t = 0.f;
        for (int k = 0; k < ba.nb; ++k) {
            if (k == mid) continue;
            const float p = a[(size_t)i + (size_t)k * pitch] * x[(long long)i + ba.off[k]];
            t = t + p;
        }
        t = b[i] - t;
        const float u = omega * t;
        const float w = u / a[(size_t)i + (size_t)mid * pitch];
        const float z = om1 * x[i];
        xx[i] = w + z;
    }
}

gm_status spmv_launch(float* out, const BandedMatrix* A, const BandPlan& p, const float* x, const float* z, int mode,
                      gm_stream_t s);

struct HostCopy {
    BandedMatrix d{};
    float* xb = nullptr;  // padded x (device)
    float* xxb = nullptr;
    float* db = nullptr;  // b (device)
    ~HostCopy() {
        if (xb) (void)hipFree(xb);
        if (xxb) (void)hipFree(xxb);
        if (db) (void)hipFree(db);
        if (d.a) gm_bm_device_free(&d);
    }
};

// Upload a host system (x padded by xpad, read at [-xpad, rows+xpad)).
static gm_status upload(HostCopy& h, float* x, const BandedMatrix* A, const float* b, int xpad, bool twin) {
    gm_status st = gm_bm_copy_to_device(&h.d, A);
    if (st != GM_OK) return st;
    const size_t len = (size_t)A->rows + 2 * (size_t)xpad;
    GM_HIP(hipMalloc(&h.xb, sizeof(float) * (len > 0 ? len : 1)));
    GM_HIP(hipMemcpy(h.xb, x - xpad, sizeof(float) * len, hipMemcpyHostToDevice));
    if (twin) {
        GM_HIP(hipMalloc(&h.xxb, sizeof(float) * (len > 0 ? len : 1)));
        GM_HIP(hipMemcpy(h.xxb, h.xb, sizeof(float) * len, hipMemcpyDeviceToDevice));
    }
    GM_HIP(hipMalloc(&h.db, sizeof(float) * (A->rows > 0 ? A->rows : 1)));
    if (A->rows > 0) GM_HIP(hipMemcpy(h.db, b, sizeof(float) * A->rows, hipMemcpyHostToDevice));
    return GM_OK;
}

static int host_mid(const BandedMatrix* A) { return A->nbands / 2; }

// True when every entry of a Levin-structured host matrix whose 5x5 neighbour
// lies outside the W x H image is zero (then the 2-D wavefront order equals the
// 1-D lexicographic order on every non-zero coupling).
static bool wrap_entries_zero(const BandedMatrix* A, int W, int H) {
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) {
            const float* band = A->a + (size_t)(5 * (dy + 2) + (dx + 2)) * A->apitch;
            for (int v = 0; v < H; ++v)
                for (int u = 0; u < W; ++u) {
                    const int uu = u + dx, vv = v + dy;
                    if (uu >= 0 && uu < W && vv >= 0 && vv < H) continue;
                    if (band[(size_t)v * W + u] != 0.f) return false;
                }
        }
    return true;
}

}  // namespace gm

using namespace gm;

extern "C" {

gm_status gm_jacobi(float* x, const BandedMatrix* A, const float* b, float omega, int pad, int iterations,
                    gm_stream_t s) {
    BandPlan p;
    gm_status st = get_plan(A, &p);
    if (st != GM_OK) return st;
    if (p.args.center < 0) return fail(GM_ERR_INVALID, "jacobi: no band with offset 0");
    if (pad < p.maxoff) return fail(GM_ERR_INVALID, "jacobi: pad %d < largest offset %d", pad, p.maxoff);
    if (!x || !b) return fail(GM_ERR_INVALID, "jacobi: null vector");
    if (iterations <= 0 || A->rows == 0) return GM_OK;
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    float* xx = nullptr;
    if ((st = gm_vec_device_malloc(&xx, A->rows, pad, pad)) != GM_OK) return st;
    hipStream_t hs = (hipStream_t)s;
    float* cur = x;
    float* nxt = xx;
    for (int it = 0; it < iterations; ++it) {
        if ((st = gm_bm_jacobi_sweep(nxt, cur, A, b, omega, s)) != GM_OK) break;
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
    if (st != GM_OK) {
        gm_vec_device_free(xx, pad);
        return st;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && cur != x) e = hipMemcpyAsync(x, cur, sizeof(float) * A->rows, hipMemcpyDeviceToDevice, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    gm_vec_device_free(xx, pad);
    if (e != hipSuccess) return hip_fail(e, "jacobi", __FILE__, __LINE__);
    return GM_OK;
}

gm_status gm_jacobi_host(float* x, const BandedMatrix* A, const float* b, size_t iterations, int xpad, float omega) {
    if (!x || !A || !A->a || !A->bands || !b) return fail(GM_ERR_INVALID, "jacobi_host: null");
    BandPlan p;
    make_plan(A->bands, A->nbands, &p);
    if (xpad < p.maxoff) return fail(GM_ERR_INVALID, "jacobi_host: xpad %d < largest offset %d", xpad, p.maxoff);
    if (iterations == 0 || A->rows == 0) return GM_OK;
    HostCopy h;
    gm_status st = upload(h, x, A, b, xpad, true);
    if (st != GM_OK) return st;
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    const int g = elementwise_grid(d, A->rows, 1);
    float* cur = h.xb + xpad;
    float* nxt = h.xxb + xpad;
    for (size_t it = 0; it < iterations; ++it) {
        hipLaunchKernelGGL(k_jacobi_host_form, dim3(g), dim3(kBlock), 0, nullptr, nxt, cur, h.d.a, h.d.apitch, p.args,
                           host_mid(A), A->rows, h.db, omega);
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
    GM_LAUNCH_CHECK();
    GM_HIP(hipMemcpy(x, cur, sizeof(float) * A->rows, hipMemcpyDeviceToHost));
    return GM_OK;
}

gm_status gm_gauss_seidel_host(float* x, const BandedMatrix* A, const float* b, size_t iterations) {
    if (!x || !A || !A->a || !A->bands || !b) return fail(GM_ERR_INVALID, "gauss_seidel_host: null");
    BandPlan p;
    make_plan(A->bands, A->nbands, &p);
    if (iterations == 0 || A->rows == 0) return GM_OK;
    if (iterations > (size_t)0x7fffffff) return fail(GM_ERR_INVALID, "gauss_seidel_host: too many iterations");
    const int xpad = p.maxoff;  // the caller's x is readable at i + off (Solve.cpp contract)
    HostCopy h;
    gm_status st = upload(h, x, A, b, xpad, false);
    if (st != GM_OK) return st;
    float* xd = h.xb + xpad;
    const int mid = host_mid(A);
    if (p.levin25 && A->rows % p.W == 0 && wrap_entries_zero(A, p.W, A->rows / p.W)) {
        const int W = p.W, H = A->rows / p.W;
        hipLaunchKernelGGL(k_gs_wavefront, dim3(1), dim3(1024), 0, nullptr, xd, h.d.a, h.d.apitch, p.args, mid, W, H,
                           h.db, (int)iterations);
    } else {
        int c = 0x7fffffff;
        for (int k = 0; k < p.args.nb; ++k) {
            const int o = p.args.off[k] < 0 ? -p.args.off[k] : p.args.off[k];
            if (o > 0 && o < c) c = o;
        }
        if (c > A->rows) c = A->rows;
        if (c < 1) c = 1;
        hipLaunchKernelGGL(k_gs_chunks, dim3(1), dim3(1024), 0, nullptr, xd, h.d.a, h.d.apitch, p.args, mid, A->rows, c,
                           h.db, (int)iterations);
    }
    GM_LAUNCH_CHECK();
    GM_HIP(hipMemcpy(x, xd, sizeof(float) * A->rows, hipMemcpyDeviceToHost));
    return GM_OK;
}

}  // extern "C"
