// Known-answer SpMV test of BandedMatrixTest.cu:23-127, written against the
// drop-in C++ API: bands {-10, 0, 10}, every in-range entry 1, x_i = i.
// Expected b_i = 2i + 10 (i < 10), 3i (10 <= i < N-10), 2i - 10 (i >= N-10).
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "BandedMatrix.h"
#include "Vector.h"

int main() {
    const int N = 1000000;
    float* hx = (float*)malloc(N * sizeof(float));
    float* hb = (float*)malloc(N * sizeof(float));
    BandedMatrix hA;
    hA.nbands = 3;
    hA.rows = hA.cols = N;
    hA.apitch = N;
    hA.a = (float*)calloc((size_t)N * 3, sizeof(float));
    hA.bands = (int*)malloc(3 * sizeof(int));
    hA.bands[0] = -10;
    hA.bands[1] = 0;
    hA.bands[2] = 10;
    for (int i = 0; i < N; ++i) {
        hx[i] = (float)i;
        if (i - 10 >= 0) hA.a[i] = 1;
        hA.a[i + N] = 1;
        if (i + 10 < N) hA.a[i + 2 * N] = 1;
    }
    float *dx, *db;
    vecCopyToDevice(&dx, hx, N, 10, 10);
    vecDeviceMalloc(&db, N);
    BandedMatrix dA;
    bmCopyToDevice(&dA, &hA);
    bmAx(db, dA, dx);
    if (gmLastStatus() != 0) {
        printf("bmAx failed: %s\n", gm_last_error());
        return 2;
    }
    hipMemcpy(hb, db, N * sizeof(float), hipMemcpyDeviceToHost);
    bool ok = true;
    for (int i = 0; i < N && ok; ++i) {
        double e = i < 10 ? 10.0 + 2 * i : (i < N - 10 ? 3.0 * i : 1999970.0 + 2 * (i - (N - 10)));
        if (fabs(hb[i] - e) > 1e-6) {
            printf("b[%d]=%.5e expected %.5e\n", i, hb[i], e);
            ok = false;
        }
    }
    bmDeviceFree(&dA);
    vecDeviceFree(db);
    vecDeviceFree(dx, 10);
    printf(ok ? "Test PASSED\n" : "Test FAILED\n");
    return ok ? 0 : 1;
}
