// Compile-only check: every reference-named wrapper resolves against the C ABI.
#include "BandedMatrix.h"
#include "CompressedMatrix.h"
#include "CoordMatrix.h"
#include "SLIC.h"
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
    CoordMatrix coo;
    cooInit(&coo, 4, 4, 8);
    int ii[2] = {0, 1}, jj[2] = {1, 2};
    float kk[2] = {1.f, 2.f};
    cooAppend(&coo, ii, jj, kk, 2);
    cooAxpy(dx, &coo, dx, dx);
    CompressedMatrix csm;
    csmInit(&csm, 4, 4, 8);
    csmAx(dx, &csm, dx);
    csmAxpy(dx, &csm, dx, dx);
    cooRelease(&coo);
    unsigned char* argb = rgbToArgb(s, 4, 4);
    int labels[16];
    (void)slicSegmentation(labels, reinterpret_cast<const unsigned int*>(argb), 4, 4, 2, 10.0);
    free(argb);
    (void)gmLastStatus();
}
