/* Matting.h -- drop-in header: the reference's Matting API names (see gpumatting.hpp). */
#ifndef GPUMATTING_MATTING_H
#define GPUMATTING_MATTING_H
#include "BandedMatrix.h"
#ifdef __cplusplus
#include "gpumatting.hpp"
#endif
#endif
