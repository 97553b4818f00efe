// Host build of the solver's float -> FP16 conversion (gm_half.h), for the
// exhaustive comparison with numpy in tests/test_half.py.
#include <stddef.h>
#include <stdint.h>

#include "gm_half.h"

extern "C" void half_convert_range(uint32_t first, size_t n, uint16_t* out) {
    for (size_t k = 0; k < n; ++k) {
        const uint32_t bits = first + (uint32_t)k;
        out[k] = gm::f32_to_f16_rne(__builtin_bit_cast(float, bits));
    }
}
