// gm_sparse.hip -- compressed-row (CSR) and coordinate (COO) sparse matrices,
// and the superpixel multilevel operators built on them: the restriction P of
// a labelling, the prolongation P^T, the Galerkin coarse operator P A P^T of a
// banded matrix, and the pairwise merging of superpixels into coarser levels.
//
// The reference's sparse layer (CompressedMatrix.cu, CoordMatrix.cu) and its
// downsampling operator (gridoperators.h) are unfinished and never called;
// this file gives the operations they name their evident meaning (the merging
// levels, host code, are in gm_merge.cpp).  Ordering
// is deterministic everywhere: COO entries are put in (row, column) order by a
// stable radix sort (hipCUB), repeats keep their input order, and every sum
// runs in a fixed order with FP64 accumulation of FP32 products.
#include <hipcub/hipcub.hpp>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gm_internal.h"

namespace gm {

// ---------------------------------------------------------------------------
// kernels

// b[r] = (acc ? b[r] + s : s) + ysign * y[r] (y optional), s = sum of row r's
// products in entry order (one wave per row: lane l takes entries l, l + 64,
// ...; FP64 lane sums, then a fixed butterfly).
__global__ __launch_bounds__(kBlock) void k_csr_spmv(const int* __restrict__ p, const int* __restrict__ j,
                                                     const float* __restrict__ k, int rows,
                                                     const float* __restrict__ x, const float* __restrict__ y,
                                                     float ysign, int acc, float* __restrict__ b) {
    const int lane = (int)threadIdx.x & 63;
    const int nw = (int)(gridDim.x * kBlock) >> 6;
    for (int r = (int)(blockIdx.x * kBlock + threadIdx.x) >> 6; r < rows; r += nw) {
        const int e0 = p[r], e1 = p[r + 1];
        double s = 0.0;
        for (int e = e0 + lane; e < e1; e += 64) s += (double)(k[e] * x[j[e]]);
        s = wave_sum(s);
        if (lane == 0) {
            float v = (float)s;
            if (acc) v = b[r] + v;
            if (y) v = v + ysign * y[r];
            b[r] = v;
        }
    }
}

// sort keys of COO entries: row * cols + column; values: the entry index
__global__ __launch_bounds__(kBlock) void k_coo_keys(const int* __restrict__ i, const int* __restrict__ j, long n,
                                                     unsigned long long cols, unsigned long long* __restrict__ key,
                                                     unsigned* __restrict__ val) {
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock) {
        key[e] = (unsigned long long)i[e] * cols + (unsigned)j[e];
        val[e] = (unsigned)e;
    }
}

__global__ __launch_bounds__(kBlock) void k_coo_gather(const int* __restrict__ j, const float* __restrict__ k,
                                                       const unsigned* __restrict__ perm, long n,
                                                       int* __restrict__ oj, float* __restrict__ ok) {
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock) {
        const unsigned s = perm[e];
        oj[e] = j[s];
        ok[e] = k[s];
    }
}

// p[r] = first sorted entry whose key is >= r * stride, r = 0 .. rows (keys ascending)
template <class K>
__global__ __launch_bounds__(kBlock) void k_rowptr(const K* __restrict__ key, long n, int rows,
                                                   unsigned long long stride, int* __restrict__ p) {
    for (long r = (long)blockIdx.x * kBlock + threadIdx.x; r <= rows; r += (long)gridDim.x * kBlock) {
        const unsigned long long t = (unsigned long long)r * stride;
        long lo = 0, hi = n;
        while (lo < hi) {
            const long mid = (lo + hi) >> 1;
            if ((unsigned long long)key[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        p[r] = (int)lo;
    }
}

__global__ __launch_bounds__(kBlock) void k_label_check(const int* __restrict__ labels, long n, int nlabels,
                                                        int* __restrict__ bad) {
    int c = 0;
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < n; u += (long)gridDim.x * kBlock)
        c += (labels[u] < 0 || labels[u] >= nlabels);
    if (c) atomicAdd(bad, c);
}

__global__ __launch_bounds__(kBlock) void k_label_keys(const int* __restrict__ labels, long n,
                                                       unsigned* __restrict__ key, unsigned* __restrict__ val) {
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < n; u += (long)gridDim.x * kBlock) {
        key[u] = (unsigned)labels[u];
        val[u] = (unsigned)u;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_ones(float* __restrict__ k, long n) {
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += (long)gridDim.x * kBlock) k[e] = 1.f;
}

__global__ __launch_bounds__(kBlock) void k_sp_prolong(float* __restrict__ e, const float* __restrict__ xc,
                                                       const int* __restrict__ labels, long n, int nlabels) {
    for (long u = (long)blockIdx.x * kBlock + threadIdx.x; u < n; u += (long)gridDim.x * kBlock) {
        const int c = labels[u];
        e[u] = (c >= 0 && c < nlabels) ? xc[c] : __builtin_nanf("");
    }
}

// Galerkin slots: slot s = i * nb + band (fine row i); key = label[i] * nl +
// label[i + off[band]], or nl * nl (sorted last) where the neighbour is
// outside the matrix or the stored entry is 0.
__global__ __launch_bounds__(kBlock) void k_gal_keys(const float* __restrict__ a, size_t apitch, BandArgs ba,
                                                     int rows, const int* __restrict__ labels, int nl,
                                                     unsigned long long* __restrict__ key,
                                                     unsigned* __restrict__ val) {
    const long n = (long)rows * ba.nb;
    const unsigned long long none = (unsigned long long)nl * nl;
    for (long s = (long)blockIdx.x * kBlock + threadIdx.x; s < n; s += (long)gridDim.x * kBlock) {
        const int i = (int)(s / ba.nb), kb = (int)(s - (long)i * ba.nb);
        const long jn = (long)i + ba.off[kb];
        unsigned long long kk = none;
        if (jn >= 0 && jn < rows && a[(size_t)kb * apitch + i] != 0.f)
            kk = (unsigned long long)labels[i] * nl + (unsigned)labels[jn];
        key[s] = kk;
        val[s] = (unsigned)s;
    }
}

// heads of runs of equal keys among the first nv sorted keys
__global__ __launch_bounds__(kBlock) void k_run_heads(const unsigned long long* __restrict__ key, long nv,
                                                      unsigned* __restrict__ head) {
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < nv; e += (long)gridDim.x * kBlock)
        head[e] = (e == 0 || key[e] != key[e - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void k_run_starts(const unsigned* __restrict__ head,
                                                       const unsigned* __restrict__ rid, long nv,
                                                       unsigned* __restrict__ start) {
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < nv; e += (long)gridDim.x * kBlock)
        if (head[e]) start[rid[e]] = (unsigned)e;
}

// one wave per run: lane l sums the run's entries l, l + 64, ... (sorted =
// slot order) in FP64, then a fixed butterfly
__global__ __launch_bounds__(kBlock) void k_gal_sum(const float* __restrict__ a, size_t apitch, int nb,
                                                    const unsigned long long* __restrict__ key,
                                                    const unsigned* __restrict__ slot,
                                                    const unsigned* __restrict__ start, long nruns, long nv, int nl,
                                                    unsigned long long* __restrict__ rkey, int* __restrict__ oj,
                                                    float* __restrict__ ok) {
    const int lane = (int)threadIdx.x & 63;
    const long nw = ((long)gridDim.x * kBlock) >> 6;
    for (long r = ((long)blockIdx.x * kBlock + threadIdx.x) >> 6; r < nruns; r += nw) {
        const long e0 = start[r], e1 = r + 1 < nruns ? (long)start[r + 1] : nv;
        double s = 0.0;
        for (long e = e0 + lane; e < e1; e += 64) {
            const unsigned sl = slot[e];
            const unsigned i = sl / (unsigned)nb, kb = sl - i * (unsigned)nb;
            s += (double)a[(size_t)kb * apitch + i];
        }
        s = wave_sum(s);
        if (lane == 0) {
            const unsigned long long kk = key[e0];
            rkey[r] = kk;
            oj[r] = (int)(kk % (unsigned long long)nl);
            ok[r] = (float)s;
        }
    }
}

// ---------------------------------------------------------------------------
// host helpers

namespace {

int bit_length(unsigned long long v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b > 0 ? b : 1;
}

int grid_for(long n) {
    DevInfo d;
    if (dev_info(&d) != GM_OK) return 1;
    return elementwise_grid(d, n, 1);
}

// device buffers freed on scope exit
struct DevBufs {
    std::vector<void*> ptrs;
    ~DevBufs() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    gm_status alloc(T** p, size_t n) {
        *p = nullptr;
        void* q = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&q, n * sizeof(T));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
        ptrs.push_back(q);
        *p = static_cast<T*>(q);
        return GM_OK;
    }
};

#define GM_TRY(x)                      \
    do {                               \
        const gm_status st_ = (x);     \
        if (st_ != GM_OK) return st_;  \
    } while (0)

// frees an output matrix that a builder allocated unless the build finished
struct OutGuard {
    CompressedMatrix* m;
    bool done = false;
    explicit OutGuard(CompressedMatrix* m_) : m(m_) {}
    ~OutGuard() {
        if (!done && m && m->p) gm_csm_free(m);
    }
};

// stable sort of (key, value) pairs over bits [0, end_bit)
template <class K>
gm_status sort_pairs(DevBufs& B, const K* kin, K* kout, const unsigned* vin, unsigned* vout, long n, int end_bit,
                     hipStream_t s) {
    size_t tmp = 0;
    GM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, (int)n, 0, end_bit, s));
    void* t = nullptr;
    GM_TRY(B.alloc(reinterpret_cast<char**>(&t), tmp));
    GM_HIP(hipcub::DeviceRadixSort::SortPairs(t, tmp, kin, kout, vin, vout, (int)n, 0, end_bit, s));
    return GM_OK;
}

gm_status check_labels(DevBufs& B, const int* labels, long n, int nlabels, hipStream_t s, const char* who) {
    int* bad = nullptr;
    GM_TRY(B.alloc(&bad, 1));
    GM_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_label_check, dim3(grid_for(n)), dim3(kBlock), 0, s, labels, n, nlabels, bad);
    GM_LAUNCH_CHECK();
    int h = 0;
    GM_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    GM_HIP(hipStreamSynchronize(s));
    if (h) return fail(GM_ERR_INVALID, "%s: %d labels outside [0, %d)", who, h, nlabels);
    return GM_OK;
}

}  // namespace

}  // namespace gm

using namespace gm;

extern "C" {

gm_status gm_csm_init(CompressedMatrix* m, int rows, int cols, size_t length) {
    if (!m || rows < 0 || cols < 0 || length > (size_t)INT32_MAX) return fail(GM_ERR_INVALID, "gm_csm_init: arguments");
    memset(m, 0, sizeof(*m));
    GM_HIP(hipMalloc(&m->p, sizeof(int) * ((size_t)rows + 1)));
    GM_HIP(hipMemset(m->p, 0, sizeof(int) * ((size_t)rows + 1)));
    if (length) {
        if (hipMalloc(&m->j, sizeof(int) * length) != hipSuccess || hipMalloc(&m->k, sizeof(float) * length) != hipSuccess) {
            gm_csm_free(m);
            return fail(GM_ERR_OOM, "gm_csm_init: %zu entries", length);
        }
    }
    m->rows = rows;
    m->cols = cols;
    m->length = length;
    return GM_OK;
}

gm_status gm_csm_free(CompressedMatrix* m) {
    if (!m) return fail(GM_ERR_INVALID, "gm_csm_free: null");
    if (m->p) (void)hipFree(m->p);
    if (m->j) (void)hipFree(m->j);
    if (m->k) (void)hipFree(m->k);
    memset(m, 0, sizeof(*m));
    return GM_OK;
}

gm_status gm_coo_init(CoordMatrix* m, int rows, int cols, size_t length) {
    if (!m || rows < 0 || cols < 0 || length > (size_t)INT32_MAX) return fail(GM_ERR_INVALID, "gm_coo_init: arguments");
    memset(m, 0, sizeof(*m));
    if (length) {
        if (hipMalloc(&m->i, sizeof(int) * length) != hipSuccess || hipMalloc(&m->j, sizeof(int) * length) != hipSuccess ||
            hipMalloc(&m->k, sizeof(float) * length) != hipSuccess) {
            gm_coo_release(m);
            return fail(GM_ERR_OOM, "gm_coo_init: %zu entries", length);
        }
    }
    m->rows = rows;
    m->cols = cols;
    m->length = length;
    return GM_OK;
}

gm_status gm_coo_release(CoordMatrix* m) {
    if (!m) return fail(GM_ERR_INVALID, "gm_coo_release: null");
    if (m->i) (void)hipFree(m->i);
    if (m->j) (void)hipFree(m->j);
    if (m->k) (void)hipFree(m->k);
    memset(m, 0, sizeof(*m));
    return GM_OK;
}

gm_status gm_coo_append(CoordMatrix* m, const int* i, const int* j, const float* k, size_t n) {
    if (!m || (n && (!i || !j || !k))) return fail(GM_ERR_INVALID, "gm_coo_append: null");
    if (m->nnz + n > m->length) return fail(GM_ERR_INVALID, "gm_coo_append: %zu + %zu entries exceed length %zu", m->nnz, n, m->length);
    for (size_t e = 0; e < n; ++e)
        if (i[e] < 0 || i[e] >= m->rows || j[e] < 0 || j[e] >= m->cols)
            return fail(GM_ERR_INVALID, "gm_coo_append: entry %zu (%d, %d) outside %d x %d", e, i[e], j[e], m->rows, m->cols);
    if (n) {
        GM_HIP(hipMemcpy(m->i + m->nnz, i, sizeof(int) * n, hipMemcpyHostToDevice));
        GM_HIP(hipMemcpy(m->j + m->nnz, j, sizeof(int) * n, hipMemcpyHostToDevice));
        GM_HIP(hipMemcpy(m->k + m->nnz, k, sizeof(float) * n, hipMemcpyHostToDevice));
    }
    m->nnz += n;
    return GM_OK;
}

gm_status gm_coo_to_csm(CompressedMatrix* out, const CoordMatrix* in, gm_stream_t stream) {
    if (!out || !in || in->rows < 0 || in->cols < 0 || in->nnz > in->length)
        return fail(GM_ERR_INVALID, "gm_coo_to_csm: arguments");
    hipStream_t s = (hipStream_t)stream;
    const long n = (long)in->nnz;
    GM_TRY(gm_csm_init(out, in->rows, in->cols, (size_t)n));
    out->nnz = (size_t)n;
    const unsigned long long cols = in->cols > 0 ? (unsigned long long)in->cols : 1ull;
    if (n == 0) return GM_OK;
    OutGuard guard(out);
    DevBufs B;
    unsigned long long *k0, *k1;
    unsigned *v0, *v1;
    GM_TRY(B.alloc(&k0, n));
    GM_TRY(B.alloc(&k1, n));
    GM_TRY(B.alloc(&v0, n));
    GM_TRY(B.alloc(&v1, n));
    hipLaunchKernelGGL(k_coo_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, in->i, in->j, n, cols, k0, v0);
    GM_LAUNCH_CHECK();
    GM_TRY(sort_pairs(B, k0, k1, v0, v1, n, bit_length((unsigned long long)in->rows * cols), s));
    hipLaunchKernelGGL(k_coo_gather, dim3(grid_for(n)), dim3(kBlock), 0, s, in->j, in->k, v1, n, out->j, out->k);
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rowptr<unsigned long long>, dim3(grid_for((long)in->rows + 1)), dim3(kBlock), 0, s, k1, n,
                       in->rows, cols, out->p);
    GM_LAUNCH_CHECK();
    GM_HIP(hipStreamSynchronize(s));
    guard.done = true;
    return GM_OK;
}

static gm_status csr_product(float* b, const CompressedMatrix* A, const float* x, const float* y, float ysign,
                             int acc, hipStream_t s) {
    if (A->rows == 0) return GM_OK;
    long waves = A->rows;
    long blocks = (waves + kWaves - 1) / kWaves;
    DevInfo d;
    GM_TRY(dev_info(&d));
    const long cap = (long)d.cus * 16;
    hipLaunchKernelGGL(k_csr_spmv, dim3((unsigned)std::min(blocks, cap)), dim3(kBlock), 0, s, A->p, A->j, A->k, A->rows,
                       x, y, ysign, acc, b);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

gm_status gm_csm_ax(float* b, const CompressedMatrix* A, const float* x, gm_stream_t stream) {
    if (!b || !A || !A->p || (A->nnz && (!x || !A->j || !A->k))) return fail(GM_ERR_INVALID, "gm_csm_ax: null");
    return csr_product(b, A, x, nullptr, 0.f, 0, (hipStream_t)stream);
}

gm_status gm_csm_axpy(float* b, const CompressedMatrix* A, const float* x, const float* y, int add, gm_stream_t stream) {
    if (!b || !A || !A->p || !y || (A->nnz && (!x || !A->j || !A->k))) return fail(GM_ERR_INVALID, "gm_csm_axpy: null");
    return csr_product(b, A, x, y, add ? 1.f : -1.f, 0, (hipStream_t)stream);
}

gm_status gm_coo_axpy(float* b, const CoordMatrix* A, const float* x, const float* y, gm_stream_t stream) {
    if (!b || !A || !y || (A->nnz && (!x || !A->i || !A->j || !A->k))) return fail(GM_ERR_INVALID, "gm_coo_axpy: null");
    CompressedMatrix C;
    GM_TRY(gm_coo_to_csm(&C, A, stream));
    gm_status st = csr_product(b, &C, x, y, 1.f, 1, (hipStream_t)stream);
    if (st == GM_OK) {
        const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    gm_csm_free(&C);
    return st;
}

gm_status gm_sp_restriction(CompressedMatrix* P, const int* labels, int npix, int nlabels, gm_stream_t stream) {
    if (!P || !labels || npix < 0 || nlabels <= 0) return fail(GM_ERR_INVALID, "gm_sp_restriction: arguments");
    hipStream_t s = (hipStream_t)stream;
    DevBufs B;
    GM_TRY(check_labels(B, labels, npix, nlabels, s, "gm_sp_restriction"));
    GM_TRY(gm_csm_init(P, nlabels, npix, (size_t)npix));
    P->nnz = (size_t)npix;
    if (npix == 0) return GM_OK;
    OutGuard guard(P);
    unsigned *k0, *k1, *v0;
    GM_TRY(B.alloc(&k0, npix));
    GM_TRY(B.alloc(&k1, npix));
    GM_TRY(B.alloc(&v0, npix));
    hipLaunchKernelGGL(k_label_keys, dim3(grid_for(npix)), dim3(kBlock), 0, s, labels, (long)npix, k0, v0);
    GM_LAUNCH_CHECK();
    // stable: each row's pixels stay ascending
    GM_TRY(sort_pairs(B, k0, k1, v0, reinterpret_cast<unsigned*>(P->j), npix, bit_length((unsigned)nlabels), s));
    hipLaunchKernelGGL(k_fill_ones, dim3(grid_for(npix)), dim3(kBlock), 0, s, P->k, (long)npix);
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rowptr<unsigned>, dim3(grid_for((long)nlabels + 1)), dim3(kBlock), 0, s, k1, (long)npix,
                       nlabels, 1ull, P->p);
    GM_LAUNCH_CHECK();
    GM_HIP(hipStreamSynchronize(s));
    guard.done = true;
    return GM_OK;
}

gm_status gm_sp_prolong(float* e, const float* xc, const int* labels, int npix, int nlabels, gm_stream_t stream) {
    if (!e || !xc || !labels || npix < 0 || nlabels <= 0) return fail(GM_ERR_INVALID, "gm_sp_prolong: arguments");
    if (npix == 0) return GM_OK;
    hipLaunchKernelGGL(k_sp_prolong, dim3(grid_for(npix)), dim3(kBlock), 0, (hipStream_t)stream, e, xc, labels,
                       (long)npix, nlabels);
    GM_LAUNCH_CHECK();
    return GM_OK;
}

gm_status gm_sp_galerkin(CompressedMatrix* Ac, const BandedMatrix* A, const int* labels, int nlabels,
                         gm_stream_t stream) {
    if (!Ac || !A || !A->a || !A->bands || !labels || A->rows <= 0 || nlabels <= 0 || nlabels > (1 << 24))
        return fail(GM_ERR_INVALID, "gm_sp_galerkin: arguments");
    BandPlan plan;
    GM_TRY(get_plan(A, &plan));
    const long n = (long)A->rows * plan.args.nb;
    if (n >= (1L << 31)) return fail(GM_ERR_INVALID, "gm_sp_galerkin: rows x bands must stay below 2^31");
    hipStream_t s = (hipStream_t)stream;
    DevBufs B;
    GM_TRY(check_labels(B, labels, A->rows, nlabels, s, "gm_sp_galerkin"));
    const unsigned long long nl = (unsigned long long)nlabels, none = nl * nl;
    unsigned long long *k0, *k1, *rkey;
    unsigned *v0, *v1, *head, *rid, *start;
    GM_TRY(B.alloc(&k0, n));
    GM_TRY(B.alloc(&k1, n));
    GM_TRY(B.alloc(&v0, n));
    GM_TRY(B.alloc(&v1, n));
    hipLaunchKernelGGL(k_gal_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, A->a, A->apitch, plan.args, A->rows, labels,
                       nlabels, k0, v0);
    GM_LAUNCH_CHECK();
    GM_TRY(sort_pairs(B, k0, k1, v0, v1, n, bit_length(none), s));
    // entries in use: keys below `none`
    int* dnv;
    GM_TRY(B.alloc(&dnv, 2));
    hipLaunchKernelGGL(k_rowptr<unsigned long long>, dim3(1), dim3(kBlock), 0, s, k1, n, 1, none, dnv);
    GM_LAUNCH_CHECK();
    int hnv[2] = {0, 0};
    GM_HIP(hipMemcpyAsync(hnv, dnv, sizeof(hnv), hipMemcpyDeviceToHost, s));
    GM_HIP(hipStreamSynchronize(s));
    const long nv = hnv[1];  // p[1] = first key >= 1 * none
    if (nv == 0) {
        GM_TRY(gm_csm_init(Ac, nlabels, nlabels, 0));
        return GM_OK;
    }
    GM_TRY(B.alloc(&head, nv));
    GM_TRY(B.alloc(&rid, nv));
    hipLaunchKernelGGL(k_run_heads, dim3(grid_for(nv)), dim3(kBlock), 0, s, k1, nv, head);
    GM_LAUNCH_CHECK();
    size_t tmp = 0;
    GM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, head, rid, (int)nv, s));
    char* t;
    GM_TRY(B.alloc(&t, tmp));
    GM_HIP(hipcub::DeviceScan::ExclusiveSum(t, tmp, head, rid, (int)nv, s));
    unsigned hl[2];
    GM_HIP(hipMemcpyAsync(&hl[0], rid + nv - 1, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    GM_HIP(hipMemcpyAsync(&hl[1], head + nv - 1, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    GM_HIP(hipStreamSynchronize(s));
    const long nruns = (long)hl[0] + hl[1];
    GM_TRY(B.alloc(&start, nruns));
    GM_TRY(B.alloc(&rkey, nruns));
    hipLaunchKernelGGL(k_run_starts, dim3(grid_for(nv)), dim3(kBlock), 0, s, head, rid, nv, start);
    GM_LAUNCH_CHECK();
    GM_TRY(gm_csm_init(Ac, nlabels, nlabels, (size_t)nruns));
    Ac->nnz = (size_t)nruns;
    OutGuard guard(Ac);
    {
        DevInfo d;
        GM_TRY(dev_info(&d));
        const long blocks = std::min((nruns + kWaves - 1) / kWaves, (long)d.cus * 16);
        hipLaunchKernelGGL(k_gal_sum, dim3((unsigned)blocks), dim3(kBlock), 0, s, A->a, A->apitch, plan.args.nb, k1,
                           v1, start, nruns, nv, nlabels, rkey, Ac->j, Ac->k);
    }
    GM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rowptr<unsigned long long>, dim3(grid_for(nl + 1)), dim3(kBlock), 0, s, rkey, nruns, nlabels,
                       nl, Ac->p);
    GM_LAUNCH_CHECK();
    GM_HIP(hipStreamSynchronize(s));
    guard.done = true;
    return GM_OK;
}

}  // extern "C"
