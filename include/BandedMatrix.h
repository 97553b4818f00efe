/* BandedMatrix.h -- the banded matrix container, C-compatible.
 *
 * Same type name, field names, field order and meaning as the reference's
 * BandedMatrix (BandedMatrix.h:23-38), so code written against the reference
 * compiles unchanged:
 *   a       band data, nbands x rows dense "column-major": entry (i, band k) is
 *           a[i + k*apitch]
 *   apitch  elements (not bytes) between consecutive bands, >= rows
 *   bands   nbands offsets: band k holds A(i, i + bands[k])
 *   nbands, rows, cols
 * Device matrices made by this library have apitch rounded up to a multiple of
 * 64 elements (256-byte aligned bands) and zero-filled padding.
 */
#ifndef GPUMATTING_BANDEDMATRIX_H
#define GPUMATTING_BANDEDMATRIX_H

#include <stddef.h>

typedef struct BandedMatrix_s
{
   float* a;
   size_t apitch;
   int* bands;
   int nbands;
   int rows;
   int cols;
} BandedMatrix;

#endif /* GPUMATTING_BANDEDMATRIX_H */
