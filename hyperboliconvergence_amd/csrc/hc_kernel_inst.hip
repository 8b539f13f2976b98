// hc_kernel_inst.hip -- one instantiation of the stage kernel, selected with
// -DHC_LAW=<0|1> -DHC_BASIS=<0|1> -DHC_FLUX=<0..4> -DHC_PREC=<64|32>.
// The FP64 objects are compiled with -ffp-contract=off (bit-faithful arithmetic),
// the mixed ones with contraction allowed.
#include "hc_launch.hpp"

#if !defined(HC_LAW) || !defined(HC_BASIS) || !defined(HC_FLUX) || !defined(HC_PREC)
#error "hc_kernel_inst.hip needs HC_LAW, HC_BASIS, HC_FLUX and HC_PREC"
#endif

#define HC_CAT2(a, b) a##b
#define HC_CAT(a, b) HC_CAT2(a, b)
#define HC_NAME4(L, B, F, P) hc_stage_L##L##_B##B##_F##F##_P##P
#define HC_NAME(L, B, F, P) HC_NAME4(L, B, F, P)
#define HC_UNAME3(L, B, P) hc_upd_L##L##_B##B##_P##P
#define HC_UNAME(L, B, P) HC_UNAME3(L, B, P)

#if HC_PREC == 64
using RF = double;
constexpr int TI = hc::TI64, TJ = hc::TJ64, NT = hc::NT64;
#else
using RF = float;
constexpr int TI = hc::TI32, TJ = hc::TJ32, NT = hc::NT32;
#endif

hipError_t HC_NAME(HC_LAW, HC_BASIS, HC_FLUX, HC_PREC)(const hc::StageArgs<RF>& a, hipStream_t st)
{
    const int tiles_i = (a.m.ni + TI - 1) / TI;
    const int tiles_j = (a.m.nj + TJ - 1) / TJ;
    const int ntiles = tiles_i * tiles_j;
    hipLaunchKernelGGL((hc::k_stage<HC_LAW, HC_BASIS, HC_FLUX, RF, TI, TJ, NT>), dim3(ntiles), dim3(NT), 0, st, a,
                       tiles_j, ntiles);
    return hipGetLastError();
}

#if HC_FLUX == 0
hipError_t HC_UNAME(HC_LAW, HC_BASIS, HC_PREC)(const hc::StageArgs<RF>& a, hipStream_t st)
{
    const long long nb = (a.N + 255) / 256;
    hipLaunchKernelGGL((hc::k_stage0_update<HC_LAW, HC_BASIS, RF>), dim3((unsigned)nb), dim3(256), 0, st, a);
    return hipGetLastError();
}
#endif
