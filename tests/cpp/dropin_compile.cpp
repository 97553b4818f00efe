// Compile-only check: every reference-named wrapper resolves against the C ABI.
#include "BandedMatrix.h"
#include "Matting.h"
#include "Solve.h"
#include "Vector.h"

struct F4 {
    float x, y, z, w;
};

void use_everything(BandedMatrix hA, float* x, float* b, unsigned char* s, F4* im) {
    BandedMatrix dA;
    bmCopyToDevice(&dA, &hA);
    float* dx;
    vecDeviceMalloc(&dx, 10, 2, 2);
    vecCopyToDevice(&dx, x, 10, 2, 2);
    bmAx(dx, dA, dx);
    bmAxpy(dx, dA, dx, dx, true);
    jacobiSweep(dx, dx, dA, b, 0.5f);
    vecAdd(dx, dx, dx, 10);
    vecSub(dx, dx, dx, 10);
    vecScaleConst(dx, dx, 2.f, 10);
    vecScale(dx, dx, dx, 10);
    innerProd(dx, dx, dx, 10);
    gradSolve(dx, dA, b, 3, 2);
    cgSolve(dx, dA, b, 2, 3, 101);
    jacobi(dx, dA, b, 2.f / 3.f, 2, 4);
    gaussSeidel_host(x, hA, b);
    jacobi_host(x, hA, b, 2, 2);
    hostLevinLaplacian(hA, b, 1e-2f, im, s, 4, 4, 4);
    vecDeviceFree(dx, 2);
    bmDeviceFree(&dA);
    (void)gmLastStatus();
}
