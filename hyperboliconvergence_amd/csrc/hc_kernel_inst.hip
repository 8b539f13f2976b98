// hc_kernel_inst.hip -- one instantiation of the stage kernel, selected with
// -DHC_LAW=<0|1> -DHC_BASIS=<0|1> -DHC_FLUX=<0..5> -DHC_PREC=<64|32>, in every tile shape
// of its precision.  The FP64 objects are compiled with -ffp-contract=off (bit-faithful
// arithmetic), the mixed ones with contraction allowed.
#include "hc_launch.hpp"

#if !defined(HC_LAW) || !defined(HC_BASIS) || !defined(HC_FLUX) || !defined(HC_PREC)
#error "hc_kernel_inst.hip needs HC_LAW, HC_BASIS, HC_FLUX and HC_PREC"
#endif

#define HC_NAME4(L, B, F, P) hc_stage_L##L##_B##B##_F##F##_P##P
#define HC_NAME(L, B, F, P) HC_NAME4(L, B, F, P)
#define HC_UNAME3(L, B, P) hc_upd_L##L##_B##B##_P##P
#define HC_UNAME(L, B, P) HC_UNAME3(L, B, P)

#if HC_PREC == 64
using RF = double;
#else
using RF = float;
#endif

// UNI: the uniform-geometry kernel of the shape; centre staging has nothing to stage there, so a
// shape that differs from another only by it shares that shape's kernel
template <int K, int UNI> static hipError_t launch_shape(const hc::LaunchArgs<RF>& L, hipStream_t st)
{
    constexpr hc::TileShape S = hc::tile_shapes(HC_PREC == 32, HC_LAW, HC_FLUX)[K];
    constexpr int CST = UNI ? 0 : S.cst;
    hipLaunchKernelGGL((hc::k_stage<HC_LAW, HC_BASIS, HC_FLUX, RF, S.ti, S.tj, S.nt, CST, S.gpf, UNI>), dim3(L.ntiles),
                       dim3(S.nt), 0, st, L);
    return hipGetLastError();
}

hipError_t HC_NAME(HC_LAW, HC_BASIS, HC_FLUX, HC_PREC)(const hc::LaunchArgs<RF>& L, int cfg, int uni, hipStream_t st)
{
    static_assert(hc::NTILECFG == 3, "one launch_shape per tile configuration");
    if (L.ntiles <= 0) return hipSuccess;
    if (uni) {
        switch (cfg) {
        case 0: return launch_shape<0, 1>(L, st);
        case 1: return launch_shape<1, 1>(L, st);
        case 2: return launch_shape<2, 1>(L, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (cfg) {
    case 0: return launch_shape<0, 0>(L, st);
    case 1: return launch_shape<1, 0>(L, st);
    case 2: return launch_shape<2, 0>(L, st);
    default: return hipErrorInvalidValue;
    }
}

#if HC_FLUX == 0
hipError_t HC_UNAME(HC_LAW, HC_BASIS, HC_PREC)(const hc::LaunchArgs<RF>& L, const hc::UpdChunk* chunks, int nchunks,
                                                hipStream_t st)
{
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL((hc::k_stage0_update<HC_LAW, HC_BASIS, RF>), dim3((unsigned)nchunks), dim3(256), 0, st, L,
                       chunks);
    return hipGetLastError();
}
#endif
