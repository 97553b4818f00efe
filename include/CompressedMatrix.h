/* CompressedMatrix.h -- compressed-row (CSR) sparse matrix, C-compatible.
 *
 * Same type name, field names, field order and meaning as the reference's
 * CompressedMatrix (CompressedMatrix.h:26-37):
 *   p       rows + 1 row pointers: row r holds entries p[r] .. p[r+1]-1
 *   j, k    column index and value of each entry (columns ascending in a row;
 *           equal columns may repeat, their values add)
 *   rows, cols
 *   nnz     entries in use; length: entries allocated for j and k
 * Matrices made by this library hold device arrays.
 */
#ifndef GPUMATTING_COMPRESSEDMATRIX_H
#define GPUMATTING_COMPRESSEDMATRIX_H

#include <stddef.h>

typedef struct CompressedMatrix_s
{
   int* p;
   int* j;
   float* k;
   int rows;
   int cols;
   size_t nnz;
   size_t length;
} CompressedMatrix;

#endif /* GPUMATTING_COMPRESSEDMATRIX_H */
