/* Solve.h -- drop-in header: the reference's Solve API names (see gpumatting.hpp). */
#ifndef GPUMATTING_SOLVE_H
#define GPUMATTING_SOLVE_H
#include "BandedMatrix.h"
#ifdef __cplusplus
#include "gpumatting.hpp"
#endif
#endif
