// gm_half.h -- float -> FP16 / BF16 round-to-nearest-even conversions in
// integer arithmetic, for host and device.
//
// The FP16 conversion is written out rather than left to the `(_Float16)f`
// cast: the gfx950 backend fuses a float product followed by that cast into
// one mixed-precision FMA (v_fma_mixlo_f16), which rounds the exact product
// once instead of rounding the float product -- different bits where the
// float product is an FP16 rounding tie -- and it does so whatever the FP
// contraction setting.  Integer arithmetic leaves nothing to fuse, and the
// same code runs on the host for the exhaustive check against numpy.
#pragma once

#ifndef GM_HD
#ifdef __HIPCC__
#define GM_HD __host__ __device__
#else
#define GM_HD
#endif
#endif

namespace gm {

// IEEE binary16, round to nearest even: normal and subnormal results,
// overflow to infinity, NaN -> quiet NaN of the same sign.
GM_HD inline unsigned short f32_to_f16_rne(float f) {
    const unsigned int x = __builtin_bit_cast(unsigned int, f);
    const unsigned int sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (unsigned short)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (ax >= 0x477ff000u) return (unsigned short)(sign | 0x7c00u);  // >= 65520 rounds to infinity
    if (ax < 0x38800000u) {                                           // below 2^-14: subnormal or zero
        if (ax < 0x33000000u) return (unsigned short)sign;           // below 2^-25 (2^-25 ties to 0 too)
        const unsigned int e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
        const unsigned int sh = 126u - e;  // 14..24: units of 2^-24
        const unsigned int q = m >> sh, rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
        return (unsigned short)(sign | (q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u)));
    }
    // rebias the exponent (127 -> 15) and drop 13 mantissa bits; a carry out
    // of the mantissa steps the exponent, as it should
    const unsigned int r = ax - 0x38000000u;
    const unsigned int q = r >> 13, rem = r & 0x1fffu;
    return (unsigned short)(sign | (q + ((rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ? 1u : 0u)));
}

// bfloat16, round to nearest even; NaN -> quiet NaN.
GM_HD inline unsigned short f32_to_bf16_rne(float f) {
    unsigned int u = __builtin_bit_cast(unsigned int, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    const unsigned int lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (unsigned short)(u >> 16);
}

}  // namespace gm
