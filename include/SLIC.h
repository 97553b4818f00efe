/* SLIC.h -- drop-in header: the reference's slicSegmentation (SLIC.h:43) over
 * the library's GPU SLIC (see gpumatting.hpp). */
#ifndef GPUMATTING_SLIC_H
#define GPUMATTING_SLIC_H
#ifdef __cplusplus
#include "gpumatting.hpp"
#endif
#endif
