// gm_blas1.hip -- BLAS-1 style vector kernels (Vector.cu:124-265 roles) and
// the per-stream reduction scratch shared by every reducing kernel.
#include <map>
#include <unordered_map>
#include <utility>

#include "gm_internal.h"

namespace gm {

// One scratch block per (device, stream): partial sums of the persistent grid
// plus the arrival counter of the last-block reduction.  Kernels on one stream
// run in order, so they may share it; different streams get different blocks.

static std::mutex g_red_mu;
static std::map<std::pair<int, gm_stream_t>, RedScratch> g_red;

gm_status red_scratch(gm_stream_t s, RedScratch* out) {
    int dev = 0;
    GM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_red_mu);
    auto key = std::make_pair(dev, s);
    auto it = g_red.find(key);
    if (it == g_red.end()) {
        RedScratch r;
        void* p = nullptr;
        const size_t bytes = sizeof(double) * 4 * kMaxGrid + 256;
        GM_HIP(hipMalloc(&p, bytes));
        GM_HIP(hipMemset(p, 0, bytes));
        r.partials = reinterpret_cast<double*>(p);
        r.counter = reinterpret_cast<unsigned int*>(reinterpret_cast<char*>(p) + sizeof(double) * 4 * kMaxGrid);
        it = g_red.emplace(key, r).first;
    }
    *out = it->second;
    return GM_OK;
}

void release_red_scratch() {
    std::lock_guard<std::mutex> g(g_red_mu);
    for (auto& kv : g_red) (void)hipFree(kv.second.partials);
    g_red.clear();
}

int elementwise_grid(const DevInfo& d, long n, int per_thread) {
    long g = (n + (long)kBlock * per_thread - 1) / ((long)kBlock * per_thread);
    long cap = (long)d.cus * 8;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

// ---------------------------------------------------------------------------

template <bool ADD>
__global__ __launch_bounds__(kBlock) void k_vec_addsub(float* r, const float* x, const float* y, int n) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        r[i] = ADD ? x[i] + y[i] : x[i] - y[i];
}

__global__ __launch_bounds__(kBlock) void k_vec_scale(float* r, const float* x, float s, const float* sp, int n) {
    const float f = sp ? *sp : s;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) r[i] = f * x[i];
}

// <x, y>: float products, double accumulation, deterministic two-stage sum.
__global__ __launch_bounds__(kBlock) void k_dot(const float* __restrict__ x, const float* __restrict__ y, int n,
                                                float* result, double* partials, unsigned int* counter) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[0] += (double)(x[i] * y[i]);
    grid_reduce<1>(v, partials, counter, [&](double* t) { *result = (float)t[0]; });
}

// Scratch of gm_inner_prod_noalloc: one FP64 partial sum per workgroup (the
// grid is at most kMaxGrid workgroups) followed by the last-block arrival
// counter, 33 KB.  Every GPU the library runs on holds its own copy until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocPartials = kMaxGrid;
constexpr size_t kNoAllocBytes = sizeof(double) * kNoAllocPartials + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

__global__ void k_clear_u32(unsigned int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

static std::mutex g_na_mu;
static std::unordered_map<int, unsigned char*> g_na_ptr;  // device -> its copy of scratch

// The current device's scratch: its address looked up once per device, and
// the arrival counter cleared then (by a kernel on stream s, ordered before
// the caller's reduction); every completed reduction leaves it at zero.
static gm_status noalloc_scratch(gm_stream_t s, RedScratch* out) {
    int dev = 0;
    GM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_na_mu);
    auto it = g_na_ptr.find(dev);
    if (it == g_na_ptr.end()) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        if (!ptr) return fail(GM_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        unsigned char* base = static_cast<unsigned char*>(ptr);
        hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, (hipStream_t)s,
                           reinterpret_cast<unsigned int*>(base + sizeof(double) * kNoAllocPartials), 1);
        GM_LAUNCH_CHECK();
        it = g_na_ptr.emplace(dev, base).first;
    }
    out->partials = reinterpret_cast<double*>(it->second);
    out->counter = reinterpret_cast<unsigned int*>(it->second + sizeof(double) * kNoAllocPartials);
    return GM_OK;
}

}  // namespace gm

using namespace gm;

static gm_status vec_check(const void* r, const void* x, const void* y, int len) {
    if (len < 0) return fail(GM_ERR_INVALID, "negative length");
    if (len > 0 && (!r || !x || !y)) return fail(GM_ERR_INVALID, "null vector");
    return GM_OK;
}

extern "C" {

gm_status gm_vec_add(float* r, const float* x, const float* y, int len, gm_stream_t s) {
    gm_status st = vec_check(r, x, y, len);
    if (st != GM_OK || len == 0) return st;
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    hipLaunchKernelGGL(k_vec_addsub<true>, dim3(elementwise_grid(d, len, 4)), dim3(kBlock), 0,
                       (hipStream_t)s, r, x, y, len);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

gm_status gm_vec_sub(float* r, const float* x, const float* y, int len, gm_stream_t s) {
    gm_status st = vec_check(r, x, y, len);
    if (st != GM_OK || len == 0) return st;
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    hipLaunchKernelGGL(k_vec_addsub<false>, dim3(elementwise_grid(d, len, 4)), dim3(kBlock), 0,
                       (hipStream_t)s, r, x, y, len);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

gm_status gm_vec_scale_const(float* r, const float* x, float sc, int len, gm_stream_t s) {
    gm_status st = vec_check(r, x, x, len);
    if (st != GM_OK || len == 0) return st;
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    hipLaunchKernelGGL(k_vec_scale, dim3(elementwise_grid(d, len, 4)), dim3(kBlock), 0, (hipStream_t)s, r,
                       x, sc, (const float*)nullptr, len);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

gm_status gm_vec_scale(float* r, const float* x, const float* sp, int len, gm_stream_t s) {
    gm_status st = vec_check(r, x, sp, len);
    if (st != GM_OK || len == 0) return st;
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    hipLaunchKernelGGL(k_vec_scale, dim3(elementwise_grid(d, len, 4)), dim3(kBlock), 0, (hipStream_t)s, r,
                       x, 0.f, sp, len);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

static int dot_grid(const DevInfo& d, int len) {
    int g = elementwise_grid(d, len > 0 ? len : 1, 8);
    return g > kMaxGrid ? kMaxGrid : g;
}

gm_status gm_inner_prod(float* result, const float* x, const float* y, int len, gm_stream_t s) {
    gm_status st = vec_check(result, x, y, len);
    if (st != GM_OK) return st;
    if (!result) return fail(GM_ERR_INVALID, "null result");
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    RedScratch rs;
    if ((st = red_scratch(s, &rs)) != GM_OK) return st;
    hipLaunchKernelGGL(k_dot, dim3(dot_grid(d, len)), dim3(kBlock), 0, (hipStream_t)s, x, y, len, result,
                       rs.partials, rs.counter);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

// gm_inner_prod with its partial sums and counter in the library's static
// device scratch: no allocation, same kernel and grid (so the same result
// bit for bit).  Calls that overlap on one GPU share that scratch and race.
gm_status gm_inner_prod_noalloc(float* result, const float* x, const float* y, int len, gm_stream_t s) {
    gm_status st = vec_check(result, x, y, len);
    if (st != GM_OK) return st;
    if (!result) return fail(GM_ERR_INVALID, "null result");
    DevInfo d;
    if ((st = dev_info(&d)) != GM_OK) return st;
    const int g = dot_grid(d, len);
    if ((size_t)g > kNoAllocPartials)
        return fail(GM_ERR_INVALID, "gm_inner_prod_noalloc: %d workgroups exceed the scratch's %zu partial sums", g,
                    kNoAllocPartials);
    RedScratch rs;
    if ((st = noalloc_scratch(s, &rs)) != GM_OK) return st;
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(kBlock), 0, (hipStream_t)s, x, y, len, result, rs.partials, rs.counter);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

}  // extern "C"
