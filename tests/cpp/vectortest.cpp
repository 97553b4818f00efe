// Known-answer inner product of VectorTest.cu:22-68 through the drop-in C++
// API: <1, 2> over 1e6 elements = 2e6; run twice to check the result is
// bitwise repeatable (the reference's atomicAdd version is not).
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "Vector.h"

int main() {
    const int N = 1000000;
    float* hx = (float*)malloc(N * sizeof(float));
    float* hy = (float*)malloc(N * sizeof(float));
    for (int i = 0; i < N; ++i) {
        hx[i] = 1.f;
        hy[i] = 2.f;
    }
    float *dx, *dy, *dr;
    vecCopyToDevice(&dx, hx, N);
    vecCopyToDevice(&dy, hy, N);
    vecDeviceMalloc(&dr, 1);
    float r1 = 0, r2 = 0;
    innerProd(dr, dx, dy, N);
    hipMemcpy(&r1, dr, sizeof(float), hipMemcpyDeviceToHost);
    innerProd(dr, dx, dy, N);
    hipMemcpy(&r2, dr, sizeof(float), hipMemcpyDeviceToHost);
    vecDeviceFree(dx);
    vecDeviceFree(dy);
    vecDeviceFree(dr);
    const bool ok = gmLastStatus() == 0 && fabs(r1 - 2e6) < 1e-5 && r1 == r2;
    printf(ok ? "Test PASSED\n" : "Test FAILED\nresult: %.2e\n", r1);
    return ok ? 0 : 1;
}
