// hc_tiles.hpp -- tile shapes and work-group sizes of the stage kernels (shared by the
// launchers and the host emulation of the kernels).
#pragma once

namespace hc {

// FP64 face arithmetic: 4x32 cells on 256 lanes (two per cell: 4 waves per SIMD within the
// LDS budget of 4 work-groups per CU); mixed: 8x32 cells on 256 lanes
constexpr int TI64 = 4, TJ64 = 32, NT64 = 256;
constexpr int TI32 = 8, TJ32 = 32, NT32 = 256;

}   // namespace hc
