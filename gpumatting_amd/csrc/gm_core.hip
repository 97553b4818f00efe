// gm_core.hip -- error state, band plans, device info, matrix/vector memory.
//
// Memory conventions follow the reference (BandedMatrix.cu:23-47,
// Vector.cu:23-46): a device matrix owns `a` and `bands`; a padded vector is
// one allocation whose returned pointer is offset by the left pad.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <unordered_map>
#include <vector>

#include "gm_internal.h"

namespace gm {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

gm_status fail(gm_status st, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return st;
}

gm_status hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? GM_ERR_OOM : GM_ERR_HIP;
}

// ---------------------------------------------------------------------------
// band plans

static std::mutex g_plan_mu;
static std::unordered_map<const int*, BandPlan> g_plans;

void make_plan(const int* hb, int nb, BandPlan* p) {
    p->args.nb = nb;
    p->args.center = -1;
    p->maxoff = 0;
    // the diagonal band: the reference's convention nbands/2 (Solve.cpp:10)
    // when its offset is 0, else the first band with offset 0 (offsets of the
    // 5x5 layout alias for W < 5)
    if (nb > 0 && nb <= kMaxBands && hb[nb / 2] == 0) p->args.center = nb / 2;
    for (int k = 0; k < nb && k < kMaxBands; ++k) {
        p->args.off[k] = hb[k];
        if (hb[k] == 0 && p->args.center < 0) p->args.center = k;
        int a = hb[k] < 0 ? -hb[k] : hb[k];
        if (a > p->maxoff) p->maxoff = a;
    }
    p->levin25 = false;
    p->W = 0;
    if (nb == 25) {
        const int W = hb[17];  // band 17 = (dy=1, dx=0)
        bool ok = W >= 5;
        for (int dy = -2; ok && dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx)
                if (hb[5 * (dy + 2) + (dx + 2)] != dy * W + dx) { ok = false; break; }
        if (ok) {
            p->levin25 = true;
            p->W = W;
        }
    }
}

void register_plan(const int* dev_bands, int nb, const int* hb) {
    BandPlan p;
    make_plan(hb, nb, &p);
    std::lock_guard<std::mutex> g(g_plan_mu);
    g_plans[dev_bands] = p;
}

void unregister_plan(const int* dev_bands) {
    std::lock_guard<std::mutex> g(g_plan_mu);
    g_plans.erase(dev_bands);
}

gm_status get_plan(const BandedMatrix* A, BandPlan* out) {
    if (!A || !A->a || !A->bands) return fail(GM_ERR_INVALID, "null matrix or band pointer");
    if (A->nbands <= 0 || A->nbands > kMaxBands)
        return fail(GM_ERR_INVALID, "nbands %d outside [1,%d]", A->nbands, kMaxBands);
    if (A->rows < 0 || (size_t)A->rows > A->apitch)
        return fail(GM_ERR_INVALID, "rows %d exceeds apitch %zu", A->rows, A->apitch);
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        auto it = g_plans.find(A->bands);
        if (it != g_plans.end() && it->second.args.nb == A->nbands) {
            *out = it->second;
            return GM_OK;
        }
    }
    // A device matrix this library did not allocate: fetch its offsets once.
    std::vector<int> hb(A->nbands);
    GM_HIP(hipMemcpy(hb.data(), A->bands, sizeof(int) * A->nbands, hipMemcpyDeviceToHost));
    register_plan(A->bands, A->nbands, hb.data());
    make_plan(hb.data(), A->nbands, out);
    return GM_OK;
}

// ---------------------------------------------------------------------------

gm_status dev_info(DevInfo* d) {
    static std::mutex mu;
    static std::unordered_map<int, int> cus;
    int dev = 0;
    GM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = cus.find(dev);
    if (it == cus.end()) {
        int n = 0;
        GM_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        it = cus.emplace(dev, n > 0 ? n : 1).first;
    }
    d->device = dev;
    d->cus = it->second;
    return GM_OK;
}

int grid_for_tiles(const DevInfo& d, long ntiles, int per_cu) {
    long g = (long)d.cus * per_cu;
    if (g > ntiles) g = ntiles;
    if (g < 1) g = 1;
    g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
    if (g > kMaxGrid) g = kMaxGrid;
    return (int)g;
}

static size_t round_pitch(size_t rows) { return (rows + 63) / 64 * 64; }

}  // namespace gm

using namespace gm;

extern "C" {

const char* gm_last_error(void) { return g_err; }

const char* gm_version(void) { return "gpumatting-amd 0.1 (gfx950)"; }

int gm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void gm_levin_bands(int W, int* bands) {
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) bands[5 * (dy + 2) + (dx + 2)] = dy * W + dx;
}

int gm_levin_pad(int W) { return ((2 * W + 2) / 32) * 32 + 32; }

gm_status gm_bm_device_alloc(BandedMatrix* dA, int rows, int nbands, const int* hb) {
    if (!dA || !hb || rows < 0 || nbands <= 0 || nbands > kMaxBands)
        return fail(GM_ERR_INVALID, "gm_bm_device_alloc: bad arguments");
    const size_t pitch = round_pitch((size_t)rows);
    float* a = nullptr;
    int* b = nullptr;
    GM_HIP(hipMalloc(&a, sizeof(float) * pitch * nbands + 256));
    hipError_t e = hipMalloc(&b, sizeof(int) * nbands);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return hip_fail(e, "hipMalloc(bands)", __FILE__, __LINE__);
    }
    GM_HIP(hipMemset(a, 0, sizeof(float) * pitch * nbands + 256));
    GM_HIP(hipMemcpy(b, hb, sizeof(int) * nbands, hipMemcpyHostToDevice));
    dA->a = a;
    dA->apitch = pitch;
    dA->bands = b;
    dA->nbands = nbands;
    dA->rows = rows;
    dA->cols = rows;
    register_plan(b, nbands, hb);
    return GM_OK;
}

gm_status gm_bm_copy_to_device(BandedMatrix* dA, const BandedMatrix* hA) {
    if (!dA || !hA || !hA->a || !hA->bands) return fail(GM_ERR_INVALID, "gm_bm_copy_to_device: null");
    if ((size_t)hA->rows > hA->apitch) return fail(GM_ERR_INVALID, "rows > apitch");
    BandedMatrix d;
    gm_status st = gm_bm_device_alloc(&d, hA->rows, hA->nbands, hA->bands);
    if (st != GM_OK) return st;
    d.cols = hA->cols;
    GM_HIP(hipMemcpy2D(d.a, d.apitch * sizeof(float), hA->a, hA->apitch * sizeof(float),
                       (size_t)hA->rows * sizeof(float), hA->nbands, hipMemcpyHostToDevice));
    *dA = d;
    return GM_OK;
}

gm_status gm_bm_copy_to_host(BandedMatrix* hA, const BandedMatrix* dA) {
    if (!dA || !hA || !hA->a || !dA->a) return fail(GM_ERR_INVALID, "gm_bm_copy_to_host: null");
    if (hA->nbands != dA->nbands || hA->rows != dA->rows || (size_t)hA->rows > hA->apitch)
        return fail(GM_ERR_INVALID, "gm_bm_copy_to_host: shape mismatch");
    GM_HIP(hipMemcpy2D(hA->a, hA->apitch * sizeof(float), dA->a, dA->apitch * sizeof(float),
                       (size_t)dA->rows * sizeof(float), dA->nbands, hipMemcpyDeviceToHost));
    return GM_OK;
}

gm_status gm_bm_device_free(BandedMatrix* dA) {
    if (!dA) return fail(GM_ERR_INVALID, "gm_bm_device_free: null");
    if (dA->bands) unregister_plan(dA->bands);
    if (dA->bands) GM_HIP(hipFree(dA->bands));
    if (dA->a) GM_HIP(hipFree(dA->a));
    dA->a = nullptr;
    dA->bands = nullptr;
    return GM_OK;
}

gm_status gm_vec_device_malloc(float** dx, int length, int left, int right) {
    if (!dx || length < 0 || left < 0 || right < 0) return fail(GM_ERR_INVALID, "gm_vec_device_malloc");
    float* p = nullptr;
    const size_t n = (size_t)length + left + right;
    GM_HIP(hipMalloc(&p, sizeof(float) * (n > 0 ? n : 1)));
    GM_HIP(hipMemset(p, 0, sizeof(float) * (n > 0 ? n : 1)));
    *dx = p + left;
    return GM_OK;
}

gm_status gm_vec_copy_to_device(float** dx, const float* hx, int length, int left, int right) {
    if (!hx && length > 0) return fail(GM_ERR_INVALID, "gm_vec_copy_to_device: null host vector");
    gm_status st = gm_vec_device_malloc(dx, length, left, right);
    if (st != GM_OK) return st;
    if (length > 0) GM_HIP(hipMemcpy(*dx, hx, sizeof(float) * length, hipMemcpyHostToDevice));
    return GM_OK;
}

gm_status gm_vec_device_free(float* dx, int left) {
    if (!dx) return GM_OK;
    GM_HIP(hipFree(dx - left));
    return GM_OK;
}

gm_status gm_vec_copy_to_host(float* hx, const float* dx, int length) {
    if (length < 0 || ((!hx || !dx) && length > 0)) return fail(GM_ERR_INVALID, "gm_vec_copy_to_host: bad argument");
    if (length > 0) GM_HIP(hipMemcpy(hx, dx, sizeof(float) * length, hipMemcpyDeviceToHost));
    return GM_OK;
}

}  // extern "C"
