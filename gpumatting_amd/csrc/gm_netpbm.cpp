// Host-side image and matrix file I/O of the matting driver: Netpbm (PGM /
// PPM) reading and writing with the reference's normalisation and
// quantisation rules, the clamped ground-truth MSE, and the A.bin banded
// matrix dump.  Plain C++ (no device code); part of libgpumatting.so so that
// the driver, the Python layer and C callers share one implementation.
//
// Formats (ppm.h:18-451 describes the same files):
//   * PGM "P5" (binary) / "P2" (ASCII), PPM "P6" / "P3"; header tokens are
//     separated by whitespace, '#' starts a comment that runs to the end of the
//     line; one whitespace byte separates maxval from binary data (or blanks
//     and a newline, as the reference's writer emits); maxval must be 1..255
//     (8-bit samples).
//   * A.bin: the matrix's bands one after another, each `rows` native-endian
//     FP32 values (entry (i, k) at float offset k * rows + i), no header, bands
//     in the matrix's own order (main.cu:334-355).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "gpumatting.h"

namespace gm {
gm_status fail(gm_status s, const char* fmt, ...);
}

namespace {

struct File {
    FILE* f = nullptr;
    explicit File(FILE* p) : f(p) {}
    ~File() {
        if (f) fclose(f);
    }
};

// Reads the next header token (skipping whitespace and comments).  *term
// receives the byte that ended the token (consumed), EOF at the end of file.
bool next_token(FILE* f, std::string& tok, int* term = nullptr) {
    tok.clear();
    if (term) *term = EOF;
    int c;
    for (;;) {
        c = fgetc(f);
        if (c == EOF) return false;
        if (c == '#') {
            while (c != '\n' && c != '\r' && c != EOF) c = fgetc(f);
            continue;
        }
        if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f') continue;
        break;
    }
    while (c != EOF && !(c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f')) {
        if (c == '#') {  // a comment right after a token ends the token
            ungetc(c, f);
            return true;
        }
        tok.push_back((char)c);
        c = fgetc(f);
    }
    if (term) *term = c;
    return true;
}

// After the maxval token of a binary file: the standard allows exactly one
// whitespace byte before the samples.  The reference's writer puts "255 \n"
// there and its reader skips to the end of that line (ppm.h:80-89, 381), so
// when maxval is followed by blanks and then a newline, the samples start
// after that newline; otherwise right after the one byte.
void skip_to_samples(FILE* f, int term) {
    if (term == '\n' || term == EOF) return;
    const long at = ftell(f);
    int c = fgetc(f);
    while (c == ' ' || c == '\t' || c == '\r') c = fgetc(f);
    if (c != '\n') fseek(f, at, SEEK_SET);
}

bool parse_int(const std::string& s, long* v) {
    if (s.empty()) return false;
    char* end = nullptr;
    const long x = strtol(s.c_str(), &end, 10);
    if (*end != '\0') return false;
    *v = x;
    return true;
}

// Reads a Netpbm file of one of the two accepted magics; returns its samples
// (ch per pixel) in row-major order.
gm_status read_pnm(const char* path, const char* bin_magic, const char* txt_magic, int ch,
                   unsigned char** out, int* w, int* h, int* maxval) {
    if (!path || !out || !w || !h) return gm::fail(GM_ERR_INVALID, "netpbm read: null argument");
    *out = nullptr;
    File f(fopen(path, "rb"));
    if (!f.f) return gm::fail(GM_ERR_INVALID, "%s: cannot open", path);
    std::string magic, ws, hs, ms;
    if (!next_token(f.f, magic)) return gm::fail(GM_ERR_INVALID, "%s: empty file", path);
    const bool binary = magic == bin_magic;
    if (!binary && magic != txt_magic)
        return gm::fail(GM_ERR_INVALID, "%s: not a %s/%s file (magic '%s')", path, bin_magic, txt_magic, magic.c_str());
    long W = 0, H = 0, M = 0;
    int term = EOF;
    if (!next_token(f.f, ws) || !next_token(f.f, hs) || !next_token(f.f, ms, &term) || !parse_int(ws, &W) ||
        !parse_int(hs, &H) || !parse_int(ms, &M))
        return gm::fail(GM_ERR_INVALID, "%s: bad header", path);
    if (W <= 0 || H <= 0 || W > (1L << 20) || H > (1L << 20) || W * H > (1L << 31) / ch)
        return gm::fail(GM_ERR_INVALID, "%s: bad size %ld x %ld", path, W, H);
    if (M < 1 || M > 255) return gm::fail(GM_ERR_INVALID, "%s: maximum value %ld is not 1..255", path, M);
    const size_t n = (size_t)W * H * ch;
    unsigned char* data = (unsigned char*)malloc(n);
    if (!data) return gm::fail(GM_ERR_OOM, "%s: out of host memory", path);
    if (binary) {
        skip_to_samples(f.f, term);
        const size_t got = fread(data, 1, n, f.f);
        if (got != n) {
            free(data);
            return gm::fail(GM_ERR_INVALID, "%s: read %zu/%zu samples", path, got, n);
        }
    } else {
        std::string t;
        for (size_t i = 0; i < n; ++i) {
            long v = 0;
            if (!next_token(f.f, t) || !parse_int(t, &v) || v < 0 || v > M) {
                free(data);
                return gm::fail(GM_ERR_INVALID, "%s: bad or missing sample %zu", path, i);
            }
            data[i] = (unsigned char)v;
        }
    }
    *out = data;
    *w = (int)W;
    *h = (int)H;
    if (maxval) *maxval = (int)M;
    return GM_OK;
}

}  // namespace

extern "C" {

void gm_free(void* p) { free(p); }

gm_status gm_pgm_read(const char* path, unsigned char** data, int* w, int* h, int* maxval) {
    return read_pnm(path, "P5", "P2", 1, data, w, h, maxval);
}

gm_status gm_ppm_read(const char* path, unsigned char** rgb, int* w, int* h, int* maxval) {
    return read_pnm(path, "P6", "P3", 3, rgb, w, h, maxval);
}

gm_status gm_pgm_write(const char* path, const unsigned char* data, int w, int h, const char* comment, int binary) {
    if (!path || !data || w <= 0 || h <= 0) return gm::fail(GM_ERR_INVALID, "pgm write: bad argument");
    File f(fopen(path, "wb"));
    if (!f.f) return gm::fail(GM_ERR_INVALID, "%s: cannot open for writing", path);
    fprintf(f.f, binary ? "P5\n" : "P2\n");
    if (comment) fprintf(f.f, "# %s \n", comment);
    fprintf(f.f, "%d %d \n255 \n", w, h);
    const size_t n = (size_t)w * h;
    if (binary) {
        if (fwrite(data, 1, n, f.f) != n) return gm::fail(GM_ERR_INVALID, "%s: short write", path);
    } else {
        for (size_t i = 0; i < n; ++i) fprintf(f.f, "%d ", (int)data[i]);
    }
    if (fflush(f.f) != 0) return gm::fail(GM_ERR_INVALID, "%s: write failed", path);
    return GM_OK;
}

void gm_quantize_u8(const float* v, int n, unsigned char* out) {
    // black = 0, white = 1, clamped; in-range values truncate 255 * v; NaN -> 0
    for (int i = 0; i < n; ++i) {
        const float x = v[i];
        out[i] = x > 1.f ? 255 : (x >= 0.f ? (unsigned char)(255.f * x) : 0);
    }
}

gm_status gm_pgm_write_float(const char* path, const float* data, int w, int h, const char* comment, int binary) {
    if (!data || w <= 0 || h <= 0) return gm::fail(GM_ERR_INVALID, "pgm write: bad argument");
    std::vector<unsigned char> q((size_t)w * h);
    gm_quantize_u8(data, w * h, q.data());
    return gm_pgm_write(path, q.data(), w, h, comment, binary);
}

gm_status gm_clamped_mse(const float* alpha, const float* gt, int n, double* mse) {
    if (!alpha || !gt || !mse || n <= 0) return gm::fail(GM_ERR_INVALID, "clamped mse: bad argument");
    // a matte value outside [0, 1] counts as its clamped value; each term is
    // formed in FP32 and summed in FP64
    double ssd = 0.0;
    for (int i = 0; i < n; ++i) {
        const float a = alpha[i], g = gt[i];
        float d;
        if (a > 1.0f) d = 1.0f - g;
        else if (a < 0.0f) d = g;
        else d = a - g;
        ssd += d * d;
    }
    *mse = ssd / n;
    return GM_OK;
}

gm_status gm_bm_dump(const BandedMatrix* A, const char* path) {
    if (!A || !A->a || !path || A->nbands <= 0 || A->rows < 0 || A->apitch < (size_t)A->rows)
        return gm::fail(GM_ERR_INVALID, "bm dump: bad matrix");
    File f(fopen(path, "wb"));
    if (!f.f) return gm::fail(GM_ERR_INVALID, "%s: cannot open for writing", path);
    for (int k = 0; k < A->nbands; ++k)
        if (fwrite(A->a + (size_t)k * A->apitch, sizeof(float), (size_t)A->rows, f.f) != (size_t)A->rows)
            return gm::fail(GM_ERR_INVALID, "%s: short write", path);
    if (fflush(f.f) != 0) return gm::fail(GM_ERR_INVALID, "%s: write failed", path);
    return GM_OK;
}

gm_status gm_bm_load(BandedMatrix* A, const char* path) {
    if (!A || !A->a || !path || A->nbands <= 0 || A->rows < 0 || A->apitch < (size_t)A->rows)
        return gm::fail(GM_ERR_INVALID, "bm load: bad matrix");
    File f(fopen(path, "rb"));
    if (!f.f) return gm::fail(GM_ERR_INVALID, "%s: cannot open", path);
    if (fseek(f.f, 0, SEEK_END) != 0) return gm::fail(GM_ERR_INVALID, "%s: cannot seek", path);
    const long size = ftell(f.f);
    const long want = (long)A->nbands * A->rows * (long)sizeof(float);
    if (size != want)
        return gm::fail(GM_ERR_INVALID, "%s: %ld bytes, expected %d bands x %d rows x 4 = %ld", path, size, A->nbands,
                        A->rows, want);
    rewind(f.f);
    for (int k = 0; k < A->nbands; ++k) {
        float* dst = A->a + (size_t)k * A->apitch;
        if (fread(dst, sizeof(float), (size_t)A->rows, f.f) != (size_t)A->rows)
            return gm::fail(GM_ERR_INVALID, "%s: short read", path);
        for (size_t i = (size_t)A->rows; i < A->apitch; ++i) dst[i] = 0.f;
    }
    return GM_OK;
}

}  // extern "C"
