// hc_adjoint_inst.hip -- launchers of the tangent and adjoint residual kernels
// (hc_adjoint.hpp), one per (law, basis, flux) in HC_FOR_EACH_ADJ; compiled with
// -ffp-contract=off so the value part matches the stage kernel bit for bit.
#include "hc_adjoint.hpp"
#include "hc_adjoint_launch.hpp"

namespace {

unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

template <int L, int B, int F> hipError_t tangent(const hc::AdjArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL((hc::k_tan_qmetric<L, B, F>), dim3(blocks(a.N)), dim3(256), 0, st, a);
    hipLaunchKernelGGL((hc::k_tan_face<L, B, F>), dim3(blocks(a.nfaces)), dim3(256), 0, st, a);
    hipLaunchKernelGGL((hc::k_tan_cell<L, B, F>), dim3(blocks(a.N)), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int L, int B, int F> hipError_t adjoint(const hc::AdjArgs& a, hipStream_t st)
{
    const long long tiles = (long long)((a.m.ni + hc::ADJ_TI - 1) / hc::ADJ_TI) * ((a.m.nj + hc::ADJ_TJ - 1) / hc::ADJ_TJ);
    hipLaunchKernelGGL((hc::k_adj_tile<L, B, F>), dim3((unsigned)tiles), dim3(hc::ADJ_NT), 0, st, a);
    hipLaunchKernelGGL((hc::k_adj_qedges<L, B, F>), dim3(blocks(a.N)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}   // namespace

#define HC_DEF_ADJ(L, B, F)                                                                                       \
    hipError_t hc_tan_L##L##_B##B##_F##F(const hc::AdjArgs& a, hipStream_t st) { return tangent<L, B, F>(a, st); } \
    hipError_t hc_adj_L##L##_B##B##_F##F(const hc::AdjArgs& a, hipStream_t st) { return adjoint<L, B, F>(a, st); }

HC_FOR_EACH_ADJ(HC_DEF_ADJ)
