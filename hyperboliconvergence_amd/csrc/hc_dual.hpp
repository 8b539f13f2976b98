// hc_dual.hpp -- forward-mode dual numbers for the tangent and adjoint kernels.
//
// Dual<N> carries a value and N tangent components.  The value is computed with
// exactly the operations of the double path (same order, same rounding), so the
// point functions of hc_point.hpp instantiated with Dual<N> return the FP64
// results in .v and their directional derivatives in .d[k].  Derivatives follow
// the elementary rules; at kinks (fabs, fmax/fmin, comparisons) the branch the
// value takes decides, which is the derivative of the discrete scheme almost
// everywhere.
#pragma once

#include "hc_point.hpp"

namespace hc {

template <int N> struct Dual {
    double v;
    double d[N];
    Dual() = default;
    HC_HD explicit Dual(double x) : v(x)
    {
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
};

template <int N> HC_HD Dual<N> make_dual(double v, int seed)
{
    Dual<N> r(v);
    if (seed >= 0 && seed < N) r.d[seed] = 1.0;
    return r;
}

template <int N> HC_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N> HC_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N> HC_HD Dual<N> operator-(const Dual<N>& a)
{
    Dual<N> r;
    r.v = -a.v;
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N> HC_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
// (a/b)' = (a' - (a/b) b') / b
template <int N> HC_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v / b.v;
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}

template <int N> HC_HD bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> HC_HD bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> HC_HD bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N> HC_HD bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }
template <int N> HC_HD bool operator==(const Dual<N>& a, const Dual<N>& b) { return a.v == b.v; }
template <int N> HC_HD bool operator!=(const Dual<N>& a, const Dual<N>& b) { return a.v != b.v; }

template <int N> HC_HD Dual<N> hc_sqrt(const Dual<N>& a)
{
    Dual<N> r;
    r.v = hc_sqrt(a.v);
    const double h = 0.5 / r.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}
template <int N> HC_HD Dual<N> hc_fabs(const Dual<N>& a)
{
    Dual<N> r;
    r.v = hc_fabs(a.v);
    const double s = a.v < 0.0 ? -1.0 : 1.0;
    for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
    return r;
}
// fmax / fmin with the double path's choice (NaN operands aside)
template <int N> HC_HD Dual<N> hc_fmax(const Dual<N>& a, const Dual<N>& b) { return hc_fmax(a.v, b.v) == a.v ? a : b; }
template <int N> HC_HD Dual<N> hc_fmin(const Dual<N>& a, const Dual<N>& b) { return hc_fmin(a.v, b.v) == a.v ? a : b; }
// a^b = exp(b log a): d = a^b (b' log a + b a'/a)
template <int N> HC_HD Dual<N> hc_pow(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = hc_pow(a.v, b.v);
    const double la = a.v > 0.0 ? ::log(a.v) : 0.0;
    const double ea = b.v / a.v;
    for (int k = 0; k < N; ++k) r.d[k] = r.v * (b.d[k] * la + ea * a.d[k]);
    return r;
}

template <int N> struct NumLimits<Dual<N>> {
    HC_HD static Dual<N> min() { return Dual<N>(DBL_MIN); }
    HC_HD static Dual<N> max() { return Dual<N>(DBL_MAX); }
};

template <int N> HC_HD Species<Dual<N>> dual_species(const Species<double>& s, int minf_seed)
{
    Species<Dual<N>> o;
    o.gamma = Dual<N>(s.gamma);
    o.minf = make_dual<N>(s.minf, minf_seed);
    o.lref = Dual<N>(s.lref);
    o.nu = Dual<N>(s.nu);
    o.pr = Dual<N>(s.pr);
    o.dt = Dual<N>(s.dt);
    o.Rg = Dual<N>(s.Rg);
    o.gamma1 = Dual<N>(s.gamma1);
    return o;
}

}   // namespace hc
