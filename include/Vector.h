/* Vector.h -- drop-in header: the reference's Vector API names (see gpumatting.hpp). */
#ifndef GPUMATTING_VECTOR_H
#define GPUMATTING_VECTOR_H
#include "BandedMatrix.h"
#ifdef __cplusplus
#include "gpumatting.hpp"
#endif
#endif
