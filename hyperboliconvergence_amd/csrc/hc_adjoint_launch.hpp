// hc_adjoint_launch.hpp -- launchers of the tangent / adjoint residual kernels
// (Euler, conserved or primitive solution basis, fluxes with a differentiable form).
#pragma once

#include "hc_adjoint.hpp"

#define HC_FOR_EACH_ADJ(X) X(1, 0, 0) X(1, 0, 1) X(1, 0, 2) X(1, 0, 3) X(1, 0, 4) X(1, 0, 5) X(1, 0, 6) X(1, 1, 4)

#define HC_DECL_ADJ(L, B, F)                                                             \
    hipError_t hc_tan_L##L##_B##B##_F##F(const hc::AdjArgs& a, hipStream_t st);          \
    hipError_t hc_adj_L##L##_B##B##_F##F(const hc::AdjArgs& a, hipStream_t st);

HC_FOR_EACH_ADJ(HC_DECL_ADJ)
