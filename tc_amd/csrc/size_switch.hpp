// size_switch.hpp -- runtime grid size -> compile-time kernel instantiation.
// The list must match tcx::kSizes in tcx_internal.hpp.
#pragma once

#define TCX_SIZE_SWITCH(n, F)                                                   \
    switch (n) {                                                                \
        case 6: F(6) break;                                                     \
        case 8: F(8) break;                                                     \
        case 10: F(10) break;                                                   \
        case 12: F(12) break;                                                   \
        case 16: F(16) break;                                                   \
        case 18: F(18) break;                                                   \
        case 20: F(20) break;                                                   \
        case 24: F(24) break;                                                   \
        case 30: F(30) break;                                                   \
        case 32: F(32) break;                                                   \
        case 36: F(36) break;                                                   \
        case 40: F(40) break;                                                   \
        case 48: F(48) break;                                                   \
        default: throw ::tcx::Error(TCX_ERR_UNSUPPORTED, "unsupported FFT grid size " + std::to_string(n)); \
    }
