// Host code under AddressSanitizer + UndefinedBehaviorSanitizer: the Netpbm
// and A.bin readers of the matting driver (gm_netpbm.cpp, which parse files
// a user hands in), the superpixel merge levels (gm_merge.cpp) and the CPU
// oracle (oracle/oracle.c), on well-formed and malformed inputs, edge sizes
// included.  Built and run by tests/test_sanitize.py; any sanitizer report
// aborts the program with a non-zero exit status.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "gpumatting.h"

// the library's error recorder (gm_core.hip in the product)
namespace gm {
static char g_last[512];
gm_status fail(gm_status s, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last, sizeof g_last, fmt, ap);
    va_end(ap);
    return s;
}
}  // namespace gm

extern "C" {
void orc_bands25(int W, int* bands);
int orc_pad(int W);
void orc_levin_scatter_f32(float* a, size_t apitch, float* b, const float* im, int nch, size_t pitch,
                           const float* lam, const float* tgt, float eps, int W, int H);
void orc_levin_scatter_u8(float* a, size_t apitch, float* b, const float* im, int nch, size_t pitch,
                          const unsigned char* scribs, float eps, float gamma, int W, int H);
void orc_levin_gather_f32(float* a, size_t apitch, float* b, float* dinv, const float* im, int nch, size_t pitch,
                          const float* lam, const float* tgt, float eps, int W, int H);
void orc_levin_gather_acc_f32(float* a, size_t apitch, float* b, float* dinv, const float* im, int nch,
                              size_t pitch, const float* lam, const float* tgt, float eps, int W, int H);
void orc_levin_f64(double* a, size_t apitch, double* b, const float* im, int nch, size_t pitch, const double* lam,
                   const double* tgt, double eps, int W, int H);
void orc_levin_dense_f64(double* L, const float* im, int nch, size_t pitch, double eps, int W, int H);
void orc_bm_ax_f32(float* y, const float* a, size_t apitch, const int* bands, int nbands, int rows, const float* x);
void orc_bm_ax_f64(double* y, const double* a, size_t apitch, const int* bands, int nbands, int rows,
                   const double* x);
void orc_cg_ref_f32(float* alpha, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                    const float* b, int pad, int iterations, int restart);
void orc_gd_ref_f32(float* alpha, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                    const float* b, int pad, int iterations);
int orc_pcg_f32(float* x, const float* a, size_t apitch, const int* bands, int nbands, int rows, const float* b,
                int pad, double tol, int maxit, int restart, int precond, double* relres);
int orc_pcg_f64(double* x, const double* a, size_t apitch, const int* bands, int nbands, int rows, const double* b,
                int pad, double tol, int maxit, double* relres);
void orc_resid_data_f64(double* out, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                        const float* lam, const float* tgt, const float* x);
void orc_gs_f32(float* x, const float* a, size_t apitch, const int* bands, int nbands, int rows, const float* b,
                size_t iterations);
void orc_jacobi_f32(float* x, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                    const float* b, size_t iterations, int xpad, float omega);
void orc_levin_backward_f64(double* grad, const float* im, int nch, size_t pitch, const double* mu,
                            const double* alpha, double eps, int W, int H);
double orc_baseline_pcg_f32(float* x, const float* a, size_t apitch, const int* bands, int nbands, int rows,
                            const float* b, int pad, int maxit, double tol, int* iters);
}

static int g_fail = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            ++g_fail;                                                   \
        }                                                               \
    } while (0)

static std::string g_dir;
static std::string path(const char* name) { return g_dir + "/" + name; }
static void write_file(const char* name, const std::string& bytes) {
    FILE* f = fopen(path(name).c_str(), "wb");
    fwrite(bytes.data(), 1, bytes.size(), f);
    fclose(f);
}

static void netpbm_cases() {
    // well-formed: binary and ASCII, comments, maxval < 255
    write_file("a.pgm", std::string("P5\n# c\n3 2\n255\n") + std::string("\x01\x02\x03\x04\x05\x06", 6));
    write_file("b.pgm", "P2 3 2 # comment\n 7\n0 1 2\n3 4 7\n");
    write_file("c.ppm", std::string("P6 2 1 255 ") + std::string("\x10\x20\x30\x40\x50\x60", 6));
    write_file("d.ppm", "P3\n2 1\n255\n1 2 3 4 5 6\n");
    unsigned char* d = nullptr;
    int w = 0, h = 0, mv = 0;
    CHECK(gm_pgm_read(path("a.pgm").c_str(), &d, &w, &h, &mv) == GM_OK && w == 3 && h == 2 && d[5] == 6);
    gm_free(d);
    CHECK(gm_pgm_read(path("b.pgm").c_str(), &d, &w, &h, &mv) == GM_OK && mv == 7 && d[5] == 7);
    gm_free(d);
    CHECK(gm_ppm_read(path("c.ppm").c_str(), &d, &w, &h, &mv) == GM_OK && w == 2 && d[5] == 0x60);
    gm_free(d);
    CHECK(gm_ppm_read(path("d.ppm").c_str(), &d, &w, &h, &mv) == GM_OK && d[5] == 6);
    gm_free(d);
    // malformed: each must fail cleanly
    const char* bad[] = {"", "P5", "P5\n", "P5 3", "P5 3 2", "P5 3 2 255", "P5 3 2 255\n\x01\x02",
                         "P5 -3 2 255\n", "P5 3 2 0\n", "P5 3 2 256\n", "P5 99999999 99999999 255\n",
                         "P2 2 1 255\n1", "P2 2 1 255\n1 x", "P2 2 1 9\n1 10", "P7 1 1 255\n\x00",
                         "P5 #\n", "P5 3 2 255 # trailing comment", "P6 1 1 255\n\x01\x02"};
    int k = 0;
    for (const char* b : bad) {
        std::string name = "bad" + std::to_string(k++) + ".pnm";
        write_file(name.c_str(), std::string(b, strlen(b)));
        d = nullptr;
        CHECK(gm_pgm_read(path(name.c_str()).c_str(), &d, &w, &h, &mv) != GM_OK || (d && (gm_free(d), true)));
        d = nullptr;
        CHECK(gm_ppm_read(path(name.c_str()).c_str(), &d, &w, &h, &mv) != GM_OK || (d && (gm_free(d), true)));
    }
    CHECK(gm_pgm_read(path("missing.pgm").c_str(), &d, &w, &h, &mv) != GM_OK);
    CHECK(gm_pgm_read(nullptr, &d, &w, &h, &mv) != GM_OK);
    // writers and the float quantiser round trip
    std::vector<float> v = {-0.5f, 0.f, 0.25f, 0.5f, 1.f, 2.f};
    CHECK(gm_pgm_write_float(path("q.pgm").c_str(), v.data(), 3, 2, "made here", 1) == GM_OK);
    CHECK(gm_pgm_read(path("q.pgm").c_str(), &d, &w, &h, &mv) == GM_OK && d[0] == 0 && d[4] == 255 && d[5] == 255);
    gm_free(d);
    std::vector<unsigned char> u8(6, 9);
    CHECK(gm_pgm_write(path("w.pgm").c_str(), u8.data(), 3, 2, nullptr, 0) == GM_OK);
    CHECK(gm_pgm_write(path("w.pgm").c_str(), u8.data(), 0, 2, nullptr, 1) != GM_OK);
    double mse = -1;
    CHECK(gm_clamped_mse(v.data(), v.data(), 6, &mse) == GM_OK && mse >= 0);
    CHECK(gm_clamped_mse(v.data(), v.data(), 0, &mse) != GM_OK);
    // A.bin dump / load, including a file of the wrong size
    const int rows = 7, nb = 3;
    std::vector<float> a(nb * 8, 1.5f), back(nb * 8, -1.f);
    std::vector<int> bands = {-1, 0, 1};
    BandedMatrix A{a.data(), 8, bands.data(), nb, rows, rows};
    BandedMatrix B{back.data(), 8, bands.data(), nb, rows, rows};
    CHECK(gm_bm_dump(&A, path("A.bin").c_str()) == GM_OK);
    CHECK(gm_bm_load(&B, path("A.bin").c_str()) == GM_OK && back[8] == 1.5f && back[7] == 0.f);
    BandedMatrix C{back.data(), 8, bands.data(), nb, rows - 1, rows - 1};
    CHECK(gm_bm_load(&C, path("A.bin").c_str()) != GM_OK);
    write_file("short.bin", "abc");
    CHECK(gm_bm_load(&B, path("short.bin").c_str()) != GM_OK);
}

static void merge_cases() {
    const int W = 9, H = 7;
    std::vector<float> im(W * H * 3);
    for (int i = 0; i < W * H * 3; ++i) im[i] = (float)((i * 37) % 11) / 10.f;
    std::vector<int> lab(W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) lab[y * W + x] = (y / 3) * 3 + x / 3;
    int nl = 9;
    CHECK(gm_sp_merge_levels(lab.data(), &nl, im.data(), 3, W, W, H, 3, 0.5f) == GM_OK && nl >= 1 && nl <= 9);
    for (int v : lab) CHECK(v >= 0 && v < nl);
    std::vector<int> one(1, 0);
    int n1 = 1;
    CHECK(gm_sp_merge_levels(one.data(), &n1, im.data(), 3, 1, 1, 1, 4, 1.f) == GM_OK && n1 == 1);
    lab[3] = 99;
    nl = 9;
    CHECK(gm_sp_merge_levels(lab.data(), &nl, im.data(), 3, W, W, H, 2, 0.5f) != GM_OK);
}

static void oracle_case(int W, int H) {
    const int N = W * H, pad = orc_pad(W);
    const size_t ap = (size_t)N;
    std::vector<float> im(N * 4), lam(N), tgt(N);
    for (int i = 0; i < N * 4; ++i) im[i] = (float)((i * 7919) % 1000) / 1000.f;
    for (int i = 0; i < N; ++i) {
        lam[i] = (i % 5 == 0) ? 100.f : 0.f;
        tgt[i] = (i % 3 == 0) ? 1.f : 0.f;
    }
    int bands[25];
    orc_bands25(W, bands);
    std::vector<float> a(25 * ap), b(N), dinv(N);
    std::vector<unsigned char> scr(N, 128);
    scr[0] = 255;
    orc_levin_scatter_f32(a.data(), ap, b.data(), im.data(), 4, W, lam.data(), tgt.data(), 1e-7f, W, H);
    orc_levin_scatter_u8(a.data(), ap, b.data(), im.data(), 4, W, scr.data(), 1e-2f, 1.f, W, H);
    orc_levin_gather_f32(a.data(), ap, b.data(), dinv.data(), im.data(), 4, W, lam.data(), tgt.data(), 1e-7f, W, H);
    orc_levin_gather_acc_f32(a.data(), ap, b.data(), dinv.data(), im.data(), 4, W, lam.data(), tgt.data(), 1e-7f, W,
                             H);
    std::vector<double> a64(25 * ap), b64(N), lam64(lam.begin(), lam.end()), tgt64(tgt.begin(), tgt.end());
    orc_levin_f64(a64.data(), ap, b64.data(), im.data(), 4, W, lam64.data(), tgt64.data(), 1e-7, W, H);
    if (N <= 64) {
        std::vector<double> dense((size_t)N * N);
        orc_levin_dense_f64(dense.data(), im.data(), 4, W, 1e-5, W, H);
    }
    std::vector<float> xp(N + 2 * pad, 0.5f), y(N);
    orc_bm_ax_f32(y.data(), a.data(), ap, bands, 25, N, xp.data() + pad);
    std::vector<double> xp64(N + 2 * pad, 0.25), y64(N);
    orc_bm_ax_f64(y64.data(), a64.data(), ap, bands, 25, N, xp64.data() + pad);
    std::vector<float> x(N + 2 * pad, 0.f);
    orc_cg_ref_f32(x.data() + pad, a.data(), ap, bands, 25, N, b.data(), pad, 12, 5);
    orc_gd_ref_f32(x.data() + pad, a.data(), ap, bands, 25, N, b.data(), pad, 6);
    double rr = 0;
    std::fill(x.begin(), x.end(), 0.f);
    orc_pcg_f32(x.data() + pad, a.data(), ap, bands, 25, N, b.data(), pad, 1e-6, 200, 7, 1, &rr);
    std::vector<double> x64(N + 2 * pad, 0.0);
    orc_pcg_f64(x64.data() + pad, a64.data(), ap, bands, 25, N, b64.data(), pad, 1e-12, 500, &rr);
    double out[2];
    std::vector<float> a0(25 * ap), b0(N), d0(N);
    orc_levin_gather_acc_f32(a0.data(), ap, b0.data(), d0.data(), im.data(), 4, W, nullptr, nullptr, 1e-7f, W, H);
    orc_resid_data_f64(out, a0.data(), ap, bands, 25, N, lam.data(), tgt.data(), x.data() + pad);
    orc_gs_f32(x.data() + pad, a.data(), ap, bands, 25, N, b.data(), 2);
    orc_jacobi_f32(x.data() + pad, a.data(), ap, bands, 25, N, b.data(), 3, pad, 2.f / 3.f);
    std::vector<double> grad(N * 3), mu(N, 0.1), al(N, 0.3);
    orc_levin_backward_f64(grad.data(), im.data(), 4, W, mu.data(), al.data(), 1e-7, W, H);
    int its = 0;
    std::fill(x.begin(), x.end(), 0.f);
    orc_baseline_pcg_f32(x.data() + pad, a.data(), ap, bands, 25, N, b.data(), pad, 50, 1e-6, &its);
    CHECK(its >= 0 && its <= 50);
}

int main(int argc, char** argv) {
    g_dir = argc > 1 ? argv[1] : ".";
    netpbm_cases();
    merge_cases();
    const int sizes[][2] = {{1, 1}, {2, 2}, {3, 3}, {1, 7}, {7, 1}, {4, 3}, {5, 8}, {17, 13}, {64, 3}, {33, 9}};
    for (auto& s : sizes) oracle_case(s[0], s[1]);
    if (g_fail) {
        fprintf(stderr, "%d checks failed\n", g_fail);
        return 1;
    }
    printf("host sanitizer run: ok\n");
    return 0;
}
