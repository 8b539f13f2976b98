// hc_launch.hpp -- launchers of the stage kernels, one per (law, basis, flux,
// precision) instantiation (hc_kernel_inst.hip is compiled once per combination);
// each launcher holds the kernels of every tile shape of its precision.
#pragma once

#include "hc_stage.hpp"
#include "hc_tiles.hpp"

#define HC_FOR_EACH_KERNEL(X) \
    X(0, 0, 0) X(0, 0, 1)     \
    X(1, 0, 0) X(1, 0, 1) X(1, 0, 2) X(1, 0, 3) X(1, 0, 4) X(1, 0, 5) X(1, 0, 6) \
    X(1, 1, 0) X(1, 1, 1) X(1, 1, 2) X(1, 1, 3) X(1, 1, 4) X(1, 1, 5) X(1, 1, 6)

#define HC_FOR_EACH_LAWBASIS(X) X(0, 0) X(1, 0) X(1, 1)

#define HC_DECL_STAGE(L, B, F)                                                                             \
    hipError_t hc_stage_L##L##_B##B##_F##F##_P64(const hc::LaunchArgs<double>&, int cfg, int uni, hipStream_t); \
    hipError_t hc_stage_L##L##_B##B##_F##F##_P32(const hc::LaunchArgs<float>&, int cfg, int uni, hipStream_t);
#define HC_DECL_UPD(L, B)                                                                                  \
    hipError_t hc_upd_L##L##_B##B##_P64(const hc::LaunchArgs<double>&, const hc::UpdChunk*, int, hipStream_t); \
    hipError_t hc_upd_L##L##_B##B##_P32(const hc::LaunchArgs<float>&, const hc::UpdChunk*, int, hipStream_t);

HC_FOR_EACH_KERNEL(HC_DECL_STAGE)
HC_FOR_EACH_LAWBASIS(HC_DECL_UPD)
