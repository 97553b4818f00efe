// matting -- closed-form matte from an image and scribbles, on one MI355X.
//
// Command line of the reference's driver (main.cu:106-332, help at :313-325):
//
//     matting <solver> <iter> <image>.ppm <scribbles>.pgm [<gt>.pgm] [options]
//
//   solver      "grad..." steepest descent, "cg..." conjugate gradient (restart
//               every 101 iterations), "cpu-jacobi..." damped Jacobi, anything
//               else Gauss-Seidel -- the first characters decide, as in the
//               reference (main.cu:137-144)
//   iter        iteration count (with --tol: the iteration cap)
//   image       binary or ASCII PPM; scribbles: PGM of the same size, 255 =
//               foreground, 0 = background, other values unconstrained
//   gt          optional ground-truth PGM: prints "Ground truth MSE: ..."
//
// The matte goes to alpha.pgm (8-bit, clamped), the Laplacian to A.bin (band
// after band, FP32), and "Image Size: W x H" to stdout; the assembly time to
// stderr.  Assembly uses window covariance regulariser 1e-2 and constraint
// weight 1 with the reference's rounding (main.cu:222, Matting.cu:39-219).
//
// Options beyond the reference's driver:
//   --tol T          Jacobi-preconditioned CG, with an FP64 master copy of
//                    alpha, until the FP64-checked relative residual is <= T
//                    (solver "cg" only; <iter> caps the iterations)
//   --precision P    fp32 (default), fp16 or bf16 operator for --tol
//   --eps E          window covariance regulariser (default 1e-2)
//   --out FILE       matte file (default alpha.pgm)
//   --no-dump        do not write A.bin
//
// Exit status: 0 on success, 1 on a bad command line, unreadable input, or a
// library error (its message goes to stderr).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "BandedMatrix.h"
#include "Matting.h"
#include "Solve.h"
#include "Vector.h"

namespace {

enum SolverKind { kGrad, kCG, kJacobiHost, kGaussSeidelHost };

void usage() {
    fprintf(stderr,
            "Usage: matting <solver> <iter> <image>.ppm <scribbles>.pgm [<gt>.pgm] [options]\n"
            "  solver    - \"grad\" (gradient descent), \"cg\" (conjugate gradient),\n"
            "              \"cpu-jacobi\" (damped Jacobi), \"cpu-gauss-seidel\"\n"
            "  iter      - number of solver iterations (iteration cap with --tol)\n"
            "  image     - RGB image (PPM)\n"
            "  scribbles - constraint image (PGM): 255 foreground, 0 background\n"
            "  gt        - ground-truth matte (PGM)\n"
            "options: --tol T  --precision fp32|fp16|bf16  --eps E  --out FILE  --no-dump\n");
    exit(0);
}

[[noreturn]] void die(const char* what) {
    fprintf(stderr, "ERROR: %s: %s\n", what, gm_last_error());
    exit(1);
}

void check(const char* what) {
    if (gmLastStatus() != GM_OK) die(what);
}

}  // namespace

int main(int argc, char* argv[]) {
    std::vector<const char*> pos;
    double tol = 0.0;
    int precision = GM_PREC_FP32;
    float eps = 1e-2f;
    const char* out = "alpha.pgm";
    bool write_dump = true;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const bool has_val = i + 1 < argc;
        if (a == "--tol" && has_val) {
            tol = atof(argv[++i]);
            if (!(tol > 0)) {
                fprintf(stderr, "ERROR: --tol must be > 0\n");
                return 1;
            }
        } else if (a == "--precision" && has_val) {
            const std::string p = argv[++i];
            if (p == "fp32") precision = GM_PREC_FP32;
            else if (p == "fp16") precision = GM_PREC_FP16;
            else if (p == "bf16") precision = GM_PREC_BF16;
            else {
                fprintf(stderr, "ERROR: --precision must be fp32, fp16 or bf16\n");
                return 1;
            }
        } else if (a == "--eps" && has_val) {
            eps = (float)atof(argv[++i]);
        } else if (a == "--out" && has_val) {
            out = argv[++i];
        } else if (a == "--no-dump") {
            write_dump = false;
        } else if (a.size() > 2 && a.compare(0, 2, "--") == 0) {
            fprintf(stderr, "ERROR: unknown option %s\n", a.c_str());
            return 1;
        } else {
            pos.push_back(argv[i]);
        }
    }
    if (pos.size() < 4) usage();

    SolverKind solver = kGaussSeidelHost;
    if (strncmp(pos[0], "grad", 4) == 0) solver = kGrad;
    else if (strncmp(pos[0], "cg", 2) == 0) solver = kCG;
    else if (strncmp(pos[0], "cpu-jacobi", 10) == 0) solver = kJacobiHost;
    if ((tol > 0 || precision != GM_PREC_FP32) && solver != kCG) {
        fprintf(stderr, "ERROR: --tol / --precision need the cg solver\n");
        return 1;
    }
    if (precision != GM_PREC_FP32 && !(tol > 0)) tol = 1e-6;
    const int iterations = atoi(pos[1]);

    int W = 0, H = 0, sW = 0, sH = 0, gW = 0, gH = 0;
    unsigned char* rgb = nullptr;
    gm_float4* im = ppmread_float4(&rgb, pos[2], &W, &H);
    if (!im) die(pos[2]);
    unsigned char* scribs = pgmread(pos[3], &sW, &sH);
    if (!scribs) die(pos[3]);
    if (sW != W || sH != H) {
        fprintf(stderr, "ERROR: scribbles not the same size as the image.\n  %d x %d vs. %d x %d\n", sW, sH, W, H);
        return 1;
    }
    float* gt = nullptr;
    if (pos.size() > 4) {
        gt = pgmread_float(pos[4], &gW, &gH);
        if (!gt) die(pos[4]);
        if (gW != W || gH != H) {
            fprintf(stderr, "ERROR: ground truth not the same size as the image.\n  %d x %d vs. %d x %d\n", gW, gH, W,
                    H);
            return 1;
        }
    }

    // host Laplacian: 25 bands, offsets dy*W + dx, unpadded pitch
    const int N = W * H;
    BandedMatrix L;
    L.rows = L.cols = N;
    L.nbands = 25;
    L.apitch = (size_t)N;
    std::vector<int> bands(25);
    gm_levin_bands(W, bands.data());
    L.bands = bands.data();
    std::vector<float> a((size_t)25 * N, 0.f), b(N), alpha(N, 0.5f);
    L.a = a.data();

    const auto t0 = std::chrono::steady_clock::now();
    hostLevinLaplacian(L, b.data(), eps, im, scribs, W, H, (size_t)W);
    check("Laplacian assembly");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fprintf(stderr, "Laplacian generation: %.2es\n", secs);
    if (write_dump) {
        dump(L);
        check("A.bin");
    }

    const int pad = gm_levin_pad(W);
    if (solver == kJacobiHost || solver == kGaussSeidelHost) {
        std::vector<float> xp((size_t)N + 2 * pad, 0.f);
        memcpy(xp.data() + pad, alpha.data(), sizeof(float) * N);
        if (solver == kJacobiHost) jacobi_host(xp.data() + pad, L, b.data(), (size_t)iterations, pad, 2.f / 3.f);
        else gaussSeidel_host(xp.data() + pad, L, b.data(), (size_t)iterations);
        check("host solver");
        memcpy(alpha.data(), xp.data() + pad, sizeof(float) * N);
    } else {
        BandedMatrix dL;
        bmCopyToDevice(&dL, &L);
        check("matrix copy");
        float *dB = nullptr, *dAlpha = nullptr;
        vecCopyToDevice(&dB, b.data(), N);
        check("vector copy");
        vecCopyToDevice(&dAlpha, alpha.data(), N, pad, pad);
        check("vector copy");
        if (solver == kGrad) {
            gradSolve(dAlpha, dL, dB, iterations, pad);
            check("gradSolve");
        } else if (tol > 0) {
            gm_solve_opts o;
            gm_solve_opts_default(&o);
            o.method = GM_METHOD_PCG;
            o.precision = precision;
            o.tol = tol;
            o.max_iter = iterations > 0 ? iterations : 1;
            // the FP64 master x keeps any tolerance reachable (an FP32 x
            // alone stalls near u * |A| |x| / |b|)
            o.refine_mode = 2;
            gm_solve_info info;
            const gm_status s = gm_solve(dAlpha, &dL, dB, pad, &o, &info, 0);
            if (s != GM_OK && s != GM_ERR_NOT_CONVERGED) die("solve");
            fprintf(stderr, "Jacobi-PCG: %d iterations, relative residual %.3e%s\n", info.iterations,
                    info.true_relres, s == GM_ERR_NOT_CONVERGED ? " (tolerance not reached)" : "");
        } else {
            cgSolve(dAlpha, dL, dB, pad, iterations, 101);
            check("cgSolve");
        }
        if (gm_vec_copy_to_host(alpha.data(), dAlpha, N) != GM_OK) die("copy back");
        vecDeviceFree(dAlpha, pad);
        vecDeviceFree(dB);
        bmDeviceFree(&dL);
    }

    printf("Image Size: %d x %d\n", W, H);
    if (gt) computeError(alpha.data(), gt, W, H);
    if (pgmwrite_float(out, W, H, alpha.data(), "", 1) != 0) die(out);

    free(gt);
    free(scribs);
    free(rgb);
    free(im);
    return 0;
}
