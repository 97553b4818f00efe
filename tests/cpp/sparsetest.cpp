// The sparse and SLIC drop-in API under the reference's names
// (CoordMatrix.cu, CompressedMatrix.cu, SLIC.h:43, ppm.h:151-165), with known
// answers: a 3x3 COO matrix with a repeated entry, b += A x + y, and SLIC on
// an image of two flat halves.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "CompressedMatrix.h"
#include "CoordMatrix.h"
#include "SLIC.h"
#include "Vector.h"

int main() {
    bool ok = true;
    // A = [[2 0 1] [0 3 0] [4 0 5]], the (0,0) entry given as 1.5 + 0.5
    int ii[6] = {2, 0, 1, 0, 2, 0};
    int jj[6] = {2, 2, 1, 0, 0, 0};
    float kk[6] = {5.f, 1.f, 3.f, 1.5f, 4.f, 0.5f};
    CoordMatrix A;
    cooInit(&A, 3, 3, 8);
    cooAppend(&A, ii, jj, kk, 6);
    float hx[3] = {1.f, 2.f, 3.f}, hy[3] = {1.f, 1.f, 1.f}, hb[3] = {10.f, 20.f, 30.f};
    float *dx, *dy, *db;
    vecCopyToDevice(&dx, hx, 3);
    vecCopyToDevice(&dy, hy, 3);
    vecCopyToDevice(&db, hb, 3);
    cooAxpy(db, &A, dx, dy);  // b += A x + y
    hipMemcpy(hb, db, sizeof(hb), hipMemcpyDeviceToHost);
    const float want[3] = {10.f + 5.f + 1.f, 20.f + 6.f + 1.f, 30.f + 19.f + 1.f};
    for (int r = 0; r < 3; ++r) ok = ok && fabsf(hb[r] - want[r]) < 1e-6f;
    if (!ok) printf("cooAxpy: %g %g %g\n", hb[0], hb[1], hb[2]);
    cooRelease(&A);
    vecDeviceFree(dx);
    vecDeviceFree(dy);
    vecDeviceFree(db);

    // two flat halves, 12 x 6, two superpixels requested
    const int w = 12, h = 6;
    unsigned char rgb[3 * w * h];
    for (int p = 0; p < w * h; ++p) {
        const bool left = (p % w) < w / 2;
        rgb[3 * p] = left ? 200 : 20;
        rgb[3 * p + 1] = left ? 30 : 60;
        rgb[3 * p + 2] = left ? 40 : 210;
    }
    unsigned char* argb = rgbToArgb(rgb, w, h);
    int labels[w * h];
    const int n = slicSegmentation(labels, reinterpret_cast<const unsigned int*>(argb), w, h, 2, 10.0);
    free(argb);
    bool sl = n == 2;
    for (int p = 0; p < w * h && sl; ++p) sl = labels[p] == ((p % w) < w / 2 ? 0 : 1);
    if (!sl) printf("slicSegmentation: %d labels\n", n);
    ok = ok && sl && gmLastStatus() == 0;
    printf(ok ? "Test PASSED\n" : "Test FAILED\n");
    return ok ? 0 : 1;
}
