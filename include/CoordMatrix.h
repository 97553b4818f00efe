/* CoordMatrix.h -- coordinate (COO) sparse matrix, C-compatible.
 *
 * Same field names, field order and meaning as the reference's CoordMatrix
 * (CoordMatrix.h:25-36): entry e is (i[e], j[e], k[e]) for e < nnz, in any
 * order, repeats adding; length entries are allocated.  Matrices made by this
 * library hold device arrays.
 */
#ifndef GPUMATTING_COORDMATRIX_H
#define GPUMATTING_COORDMATRIX_H

#include <stddef.h>

typedef struct CoordMatrix_s
{
   int* i;
   int* j;
   float* k;
   int rows;
   int cols;
   size_t nnz;
   size_t length;
} CoordMatrix;

#endif /* GPUMATTING_COORDMATRIX_H */
