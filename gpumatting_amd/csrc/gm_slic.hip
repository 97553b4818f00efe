// gm_slic.hip -- SLIC superpixels (the reference's slicSegmentation, SLIC.h:43,
// whose body is third-party code used as-is; this is a re-derivation from the
// published algorithm: Achanta, Shaji, Smith, Lucchi, Fua, Suesstrunk, "SLIC
// Superpixels Compared to State-of-the-Art Superpixel Methods", IEEE TPAMI
// 34(11), 2012, Algorithm 1).
//
//   1. RGB in [0, 1] -> CIELAB (sRGB transfer curve, D65 white; IEC 61966-2-1
//      and CIE 15 formulas), evaluated in FP64 and rounded to the nearest
//      multiple of 2^-16.  |L|, |a|, |b| < 128, so these values are exact in
//      float, and every FP64 sum of fewer than 2^29 of them is exact in any
//      order: the seed means of step 3 do not depend on the summation order.
//   2. Seeds on a grid of step S = round(sqrt(W H / K)): nsx x nsy cells
//      (nsx = round(W / S), nsy = round(H / S), at least 1), one seed at each
//      cell centre, moved to the lowest gradient position of its 3 x 3
//      neighbourhood (G = |lab(x+1,y) - lab(x-1,y)|^2 + |lab(x,y+1) -
//      lab(x,y-1)|^2, positions with all four neighbours inside the image;
//      ties to the first in raster order).
//   3. `iters` rounds of
//        assignment: each pixel takes, among the seeds of the 5 x 5 cells
//        around its own cell whose window |x - X| <= S, |y - Y| <= S holds it,
//        the one of least D = dc^2 + ds^2 m^2 / S^2 (ties: lower seed index);
//        a pixel no window holds keeps its label (initially its own cell's);
//        update: each seed moves to the mean (l, a, b, x, y) of the pixels it
//        took this round (one workgroup per seed over its window, FP64 sums in
//        a fixed order); a seed that took none stays.
//   4. Connectivity: every 4-connected piece of a label is a superpixel; a
//      piece smaller than max(1, S^2 / 4) pixels joins the superpixel of the
//      pixel left of its first pixel in raster order (or else above it; the
//      first piece of the image always stays).  Labels are numbered from 0 in
//      raster order of first appearance.
//
// Every float operation of steps 2-3 is a single IEEE operation (no FMA
// contraction), so from the same Lab planes the labels and seeds are exactly
// those of a float32 evaluation in the order written here.
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <string.h>

#include "gm_internal.h"

#pragma clang fp contract(off)

namespace gm {

struct SlicGrid {
    int W, H, S, nsx, nsy;
    float wd;  // m^2 / S^2
};

__device__ __forceinline__ float lab_grid(double v) { return (float)(rint(v * 65536.0) / 65536.0); }

__global__ __launch_bounds__(kBlock) void k_rgb_to_lab(const float* __restrict__ im, int nch, size_t pitch, int W,
                                                       int H, float* __restrict__ lab) {
    const long N = (long)W * H;
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock) {
        const int y = (int)(u / W), x = (int)(u - (long)y * W);
        const float* px = im + ((size_t)y * pitch + x) * nch;
        double c[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double v = (double)px[q];
            c[q] = v <= 0.04045 ? v / 12.92 : pow((v + 0.055) / 1.055, 2.4);
        }
        // linear sRGB -> XYZ (D65), divided by the white point
        const double X = (0.4124564 * c[0] + 0.3575761 * c[1] + 0.1804375 * c[2]) / 0.95047;
        const double Y = 0.2126729 * c[0] + 0.7151522 * c[1] + 0.0721750 * c[2];
        const double Z = (0.0193339 * c[0] + 0.1191920 * c[1] + 0.9503041 * c[2]) / 1.08883;
        auto f = [](double t) { return t > 0.008856 ? cbrt(t) : 7.787 * t + 16.0 / 116.0; };
        const double fx = f(X), fy = f(Y), fz = f(Z);
        lab[u] = lab_grid(116.0 * fy - 16.0);
        lab[N + u] = lab_grid(500.0 * (fx - fy));
        lab[2 * N + u] = lab_grid(200.0 * (fy - fz));
    }
}

__device__ __forceinline__ float lab_grad(const float* __restrict__ lab, long N, int W, int x, int y) {
    float g = 0.f;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float* L = lab + q * N;
        const float dx = L[(long)y * W + x + 1] - L[(long)y * W + x - 1];
        const float dy = L[(long)(y + 1) * W + x] - L[(long)(y - 1) * W + x];
        g += dx * dx + dy * dy;
    }
    return g;
}

// seeds: 5 planes of nseed floats (l, a, b, x, y)
__global__ __launch_bounds__(kBlock) void k_slic_seeds(const float* __restrict__ lab, SlicGrid g,
                                                       float* __restrict__ seeds) {
    const int ns = g.nsx * g.nsy;
    const long N = (long)g.W * g.H;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < ns; k += gridDim.x * kBlock) {
        const int iy = k / g.nsx, ix = k - iy * g.nsx;
        int sx = (int)((ix + 0.5f) * (float)g.W / (float)g.nsx);
        int sy = (int)((iy + 0.5f) * (float)g.H / (float)g.nsy);
        sx = min(max(sx, 0), g.W - 1);
        sy = min(max(sy, 0), g.H - 1);
        int bx = sx, by = sy;
        float best = INFINITY;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = sx + dx, y = sy + dy;
                if (x < 1 || x > g.W - 2 || y < 1 || y > g.H - 2) continue;
                const float gr = lab_grad(lab, N, g.W, x, y);
                if (gr < best) {
                    best = gr;
                    bx = x;
                    by = y;
                }
            }
        const long u = (long)by * g.W + bx;
        seeds[k] = lab[u];
        seeds[ns + k] = lab[N + u];
        seeds[2 * ns + k] = lab[2 * N + u];
        seeds[3 * ns + k] = (float)bx;
        seeds[4 * ns + k] = (float)by;
    }
}

__device__ __forceinline__ int cell_of(int v, int n, int extent) { return min(n - 1, (int)(((long)v * n) / extent)); }

__global__ __launch_bounds__(kBlock) void k_slic_init_labels(SlicGrid g, int* __restrict__ labels) {
    const long N = (long)g.W * g.H;
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock) {
        const int y = (int)(u / g.W), x = (int)(u - (long)y * g.W);
        labels[u] = cell_of(y, g.nsy, g.H) * g.nsx + cell_of(x, g.nsx, g.W);
    }
}

__global__ __launch_bounds__(kBlock) void k_slic_assign(const float* __restrict__ lab, SlicGrid g,
                                                        const float* __restrict__ seeds, int* __restrict__ labels) {
    const long N = (long)g.W * g.H;
    const int ns = g.nsx * g.nsy;
    const float S = (float)g.S;
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock) {
        const int y = (int)(u / g.W), x = (int)(u - (long)y * g.W);
        const int cx = cell_of(x, g.nsx, g.W), cy = cell_of(y, g.nsy, g.H);
        const float l = lab[u], a = lab[N + u], b = lab[2 * N + u];
        const float fx = (float)x, fy = (float)y;
        float best = INFINITY;
        int bk = -1;
        for (int jy = max(cy - 2, 0); jy <= min(cy + 2, g.nsy - 1); ++jy)
            for (int jx = max(cx - 2, 0); jx <= min(cx + 2, g.nsx - 1); ++jx) {
                const int k = jy * g.nsx + jx;
                const float ex = fx - seeds[3 * ns + k], ey = fy - seeds[4 * ns + k];
                if (fabsf(ex) > S || fabsf(ey) > S) continue;
                const float dl = l - seeds[k], da = a - seeds[ns + k], db = b - seeds[2 * ns + k];
                const float dc = dl * dl + da * da + db * db;
                const float ds = ex * ex + ey * ey;
                const float D = dc + ds * g.wd;
                if (D < best) {
                    best = D;
                    bk = k;
                }
            }
        if (bk >= 0) labels[u] = bk;
    }
}

// one workgroup per seed: mean of the pixels it took (label k inside its window)
__global__ __launch_bounds__(kBlock) void k_slic_update(const float* __restrict__ lab, SlicGrid g,
                                                        const int* __restrict__ labels,
                                                        const float* __restrict__ sin, float* __restrict__ sout) {
    const long N = (long)g.W * g.H;
    const int ns = g.nsx * g.nsy;
    const float S = (float)g.S;
    __shared__ double red[6][kWaves];
    for (int k = blockIdx.x; k < ns; k += gridDim.x) {
        const float X = sin[3 * ns + k], Y = sin[4 * ns + k];
        const int x0 = max(0, (int)floorf(X - S)), x1 = min(g.W - 1, (int)ceilf(X + S));
        const int y0 = max(0, (int)floorf(Y - S)), y1 = min(g.H - 1, (int)ceilf(Y + S));
        const int bw = x1 - x0 + 1, nbox = bw * (y1 - y0 + 1);
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int e = threadIdx.x; e < nbox; e += kBlock) {
            const int y = y0 + e / bw, x = x0 + e % bw;
            const long u = (long)y * g.W + x;
            if (labels[u] != k || fabsf((float)x - X) > S || fabsf((float)y - Y) > S) continue;
            acc[0] += lab[u];
            acc[1] += lab[N + u];
            acc[2] += lab[2 * N + u];
            acc[3] += x;
            acc[4] += y;
            acc[5] += 1.0;
        }
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double s = wave_sum(acc[q]);
            if (lane == 0) red[q][wid] = s;
        }
        __syncthreads();
        if (threadIdx.x < 5) {
            const int q = threadIdx.x;
            double t = 0, c = 0;
            for (int w = 0; w < kWaves; ++w) {
                t += red[q][w];
                c += red[5][w];
            }
            sout[q * ns + k] = c > 0 ? (float)(t / c) : sin[q * ns + k];
        }
        __syncthreads();  // red is reused by the next seed
    }
}

// Connectivity.  Union-find over the 4-neighbour edges that join equal
// labels: a union links the larger root under the smaller (atomicMin; a link
// that lost a race continues with the parent it found, the merge of Playne &
// Hawick, "A new algorithm for parallel connected-component labelling on GPUs",
// IEEE TPDS 2018), so every component's root is its first pixel in raster
// order whatever the interleaving.  Then sizes, the piece a small component
// joins (the component left of, or else above, its root), chains of joins
// resolved, and survivors numbered by a scan over their roots.
__device__ __forceinline__ int cc_parent(const int* parent, int u) {
    return __hip_atomic_load(parent + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int cc_find(const int* parent, int u) {
    for (int p = cc_parent(parent, u); p != u; p = cc_parent(parent, u)) u = p;
    return u;
}
__device__ void cc_unite(int* parent, int a, int b) {
    for (;;) {
        a = cc_find(parent, a);
        b = cc_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(parent + b, a);  // parent[v] <= v always
        if (old == b) return;
        b = old;
    }
}

__global__ __launch_bounds__(kBlock) void k_cc_init(int* __restrict__ parent, int* __restrict__ size, long N) {
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock) {
        parent[u] = (int)u;
        size[u] = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_cc_union(const int* __restrict__ labels, int W, long N, int* parent) {
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock) {
        const int x = (int)(u % W);
        if (x > 0 && labels[u - 1] == labels[u]) cc_unite(parent, (int)u, (int)u - 1);
        if (u >= W && labels[u - W] == labels[u]) cc_unite(parent, (int)u, (int)(u - W));
    }
}

__global__ __launch_bounds__(kBlock) void k_cc_flatten(int* parent, int* __restrict__ size, long N) {
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock) {
        const int r = cc_find(parent, (int)u);
        parent[u] = r;
        atomicAdd(size + r, 1);
    }
}

// roots: the root whose component a small one joins (itself when it stays)
__global__ __launch_bounds__(kBlock) void k_cc_target(const int* __restrict__ parent, const int* __restrict__ size,
                                                      int W, long N, int minsize, int* __restrict__ tgt) {
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock) {
        if (parent[u] != u) continue;
        int t = (int)u;
        if (u > 0 && size[u] < minsize) t = parent[(u % W) > 0 ? u - 1 : u - W];
        tgt[u] = t;
    }
}

// chains of joins (each step goes to a smaller root), survivors flagged
__global__ __launch_bounds__(kBlock) void k_cc_resolve(const int* __restrict__ parent, const int* __restrict__ tgt,
                                                       long N, int* __restrict__ fin, unsigned* __restrict__ flag) {
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock) {
        unsigned f = 0;
        if (parent[u] == u) {
            int r = (int)u;
            while (tgt[r] != r) r = tgt[r];
            fin[u] = r;
            f = tgt[u] == u;
        }
        flag[u] = f;
    }
}

__global__ __launch_bounds__(kBlock) void k_cc_label(const int* __restrict__ parent, const int* __restrict__ fin,
                                                     const unsigned* __restrict__ num, long N,
                                                     int* __restrict__ labels) {
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long)gridDim.x * kBlock)
        labels[u] = (int)num[fin[parent[u]]];
}

}  // namespace gm

using namespace gm;

static SlicGrid slic_grid(int W, int H, int K, float compactness) {
    SlicGrid g;
    const long N = (long)W * H;
    g.W = W;
    g.H = H;
    g.S = (int)(sqrt((double)N / K) + 0.5);
    if (g.S < 1) g.S = 1;
    g.nsx = (int)((double)W / g.S + 0.5);
    g.nsy = (int)((double)H / g.S + 0.5);
    if (g.nsx < 1) g.nsx = 1;
    if (g.nsy < 1) g.nsy = 1;
    g.wd = compactness * compactness / (float)(g.S * g.S);
    return g;
}

// steps 2-3 from Lab planes: raw labels (seed indices), final seeds into sout
// when given (5 planes of ns floats)
static gm_status slic_raw(int* labels, float* sout, const float* lab, const SlicGrid& g, int iters, hipStream_t s) {
    const long N = (long)g.W * g.H;
    const int ns = g.nsx * g.nsy;
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    float *s0 = nullptr, *s1 = nullptr;
    if (hipMalloc(&s0, sizeof(float) * 5 * ns) != hipSuccess || hipMalloc(&s1, sizeof(float) * 5 * ns) != hipSuccess) {
        if (s0) (void)hipFree(s0);
        return fail(GM_ERR_OOM, "gm_slic: %d seeds", ns);
    }
    const int grid = elementwise_grid(d, N, 1);
    hipLaunchKernelGGL(k_slic_seeds, dim3((ns + kBlock - 1) / kBlock), dim3(kBlock), 0, s, lab, g, s0);
    hipLaunchKernelGGL(k_slic_init_labels, dim3(grid), dim3(kBlock), 0, s, g, labels);
    const int ugrid = ns < d.cus * 8 ? ns : d.cus * 8;
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_slic_assign, dim3(grid), dim3(kBlock), 0, s, lab, g, s0, labels);
        hipLaunchKernelGGL(k_slic_update, dim3(ugrid), dim3(kBlock), 0, s, lab, g, labels, s0, s1);
        std::swap(s0, s1);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && sout) e = hipMemcpyAsync(sout, s0, sizeof(float) * 5 * ns, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(s0);
    (void)hipFree(s1);
    if (e != hipSuccess) return hip_fail(e, "gm_slic", __FILE__, __LINE__);
    return GM_OK;
}

// step 4 in place
static gm_status slic_connect(int* labels, int* nlabels, int W, int H, int minsize, hipStream_t s) {
    const long N = (long)W * H;
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    const int grid = elementwise_grid(d, N, 1);
    // parent, size, tgt, fin as ints; flag, num as unsigned
    int* cc = nullptr;
    void* tmp = nullptr;
    size_t ntmp = 0;
    hipError_t e = hipMalloc(&cc, sizeof(int) * 6 * (size_t)N);
    int *parent = cc, *size = cc + N, *tgt = cc + 2 * N, *fin = cc + 3 * N;
    unsigned *flag = reinterpret_cast<unsigned*>(cc + 4 * N), *num = reinterpret_cast<unsigned*>(cc + 5 * N);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, ntmp, flag, num, (int)N, s);
    if (e == hipSuccess) e = hipMalloc(&tmp, ntmp ? ntmp : 1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cc_init, dim3(grid), dim3(kBlock), 0, s, parent, size, N);
        hipLaunchKernelGGL(k_cc_union, dim3(grid), dim3(kBlock), 0, s, labels, W, N, parent);
        hipLaunchKernelGGL(k_cc_flatten, dim3(grid), dim3(kBlock), 0, s, parent, size, N);
        hipLaunchKernelGGL(k_cc_target, dim3(grid), dim3(kBlock), 0, s, parent, size, W, N, minsize, tgt);
        hipLaunchKernelGGL(k_cc_resolve, dim3(grid), dim3(kBlock), 0, s, parent, tgt, N, fin, flag);
        e = hipcub::DeviceScan::ExclusiveSum(tmp, ntmp, flag, num, (int)N, s);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cc_label, dim3(grid), dim3(kBlock), 0, s, parent, fin, num, N, labels);
        e = hipGetLastError();
    }
    unsigned last[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(&last[0], num + N - 1, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&last[1], flag + N - 1, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (cc) (void)hipFree(cc);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "gm_slic_connect", __FILE__, __LINE__);
    *nlabels = (int)(last[0] + last[1]);
    return GM_OK;
}

static bool slic_size_ok(int W, int H) { return W > 0 && H > 0 && (long)W * H < (1L << 31); }

extern "C" {

gm_status gm_slic_grid(int W, int H, int K, int* S, int* nsx, int* nsy) {
    if (!slic_size_ok(W, H) || K <= 0) return fail(GM_ERR_INVALID, "gm_slic_grid: arguments");
    const SlicGrid g = slic_grid(W, H, K, 1.f);
    if (S) *S = g.S;
    if (nsx) *nsx = g.nsx;
    if (nsy) *nsy = g.nsy;
    return GM_OK;
}

gm_status gm_slic_lab(float* lab, const float* im, int nch, size_t pitch, int W, int H, gm_stream_t stream) {
    if (!lab || !im || nch < 3 || !slic_size_ok(W, H) || pitch < (size_t)W)
        return fail(GM_ERR_INVALID, "gm_slic_lab: arguments");
    DevInfo d;
    gm_status st = dev_info(&d);
    if (st != GM_OK) return st;
    hipLaunchKernelGGL(k_rgb_to_lab, dim3(elementwise_grid(d, (long)W * H, 1)), dim3(kBlock), 0, (hipStream_t)stream,
                       im, nch, pitch, W, H, lab);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

gm_status gm_slic_raw(int* labels, float* seeds, int maxseeds, const float* lab, int W, int H, int K,
                      float compactness, int iters, gm_stream_t stream) {
    if (!labels || !lab || !slic_size_ok(W, H) || K <= 0 || !(compactness > 0.f) || iters < 0)
        return fail(GM_ERR_INVALID, "gm_slic_raw: arguments");
    const SlicGrid g = slic_grid(W, H, K, compactness);
    if (seeds && maxseeds < g.nsx * g.nsy)
        return fail(GM_ERR_INVALID, "gm_slic_raw: %d seeds do not fit %d", g.nsx * g.nsy, maxseeds);
    return slic_raw(labels, seeds, lab, g, iters, (hipStream_t)stream);
}

gm_status gm_slic_connect(int* labels, int* nlabels, int W, int H, int minsize, gm_stream_t stream) {
    if (!labels || !nlabels || !slic_size_ok(W, H) || minsize < 1)
        return fail(GM_ERR_INVALID, "gm_slic_connect: arguments");
    return slic_connect(labels, nlabels, W, H, minsize, (hipStream_t)stream);
}

gm_status gm_slic(int* labels, int* nlabels, const float* im, int nch, size_t pitch, int W, int H, int K,
                  float compactness, int iters, gm_stream_t stream) {
    if (!labels || !nlabels || !im || nch < 3 || !slic_size_ok(W, H) || pitch < (size_t)W || K <= 0 ||
        !(compactness > 0.f) || iters < 0)
        return fail(GM_ERR_INVALID, "gm_slic: arguments");
    hipStream_t s = (hipStream_t)stream;
    const long N = (long)W * H;
    const SlicGrid g = slic_grid(W, H, K, compactness);
    float* lab = nullptr;
    if (hipMalloc(&lab, sizeof(float) * 3 * N) != hipSuccess) return fail(GM_ERR_OOM, "gm_slic: %dx%d", W, H);
    gm_status st = gm_slic_lab(lab, im, nch, pitch, W, H, stream);
    if (st == GM_OK) st = slic_raw(labels, nullptr, lab, g, iters, s);
    (void)hipFree(lab);
    if (st != GM_OK) return st;
    return slic_connect(labels, nlabels, W, H, g.S * g.S / 4 > 1 ? g.S * g.S / 4 : 1, s);
}

}  // extern "C"
