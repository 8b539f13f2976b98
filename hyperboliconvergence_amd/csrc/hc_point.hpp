// hc_point.hpp -- point-wise physics of the structured-mesh finite-volume path,
// shared by the HIP kernels and the host-side C++20 drop-in API.
//
// Everything here is __host__ __device__ and templated on the arithmetic type
// R (double for the bit-faithful path, float for the mixed path's face work).
// Expressions keep the reference's evaluation order operation for operation
// (left-associative, no re-association, no fused multiply-add in the FP64
// translation unit) so that FP64 results agree with the reference's arithmetic
// to the last bit wherever the operations are correctly rounded.
//
// Reference files are cited relative to cpp/inc/ of lacrymose/hyperboliConvergence.
#pragma once

#include <cmath>
#include <cfloat>

#if defined(__HIPCC__)
#define HC_HD __host__ __device__ inline
#else
#define HC_HD inline
#endif

namespace hc {

enum Law { LAW_SCALAR = 0, LAW_EULER = 1 };
enum Basis { BASIS_CONSERVED = 0, BASIS_PRIMITIVE = 1 };
enum FluxId {
    FLUX_CENTRAL = 0, FLUX_RUSANOV = 1, FLUX_ROE = 2, FLUX_SLAU = 3, FLUX_AUSMPUP = 4, FLUX_ROE_FACEFRAME = 5,
    FLUX_ROE_UNPREC_WS = 6
};
enum Scale { SCALE_CONVECTIVE = 0, SCALE_ACOUSTIC = 1, SCALE_ADAPTIVE = 2 };
enum LimiterId {
    LIM_NOLIMIT1 = 0, LIM_NOLIMIT2 = 1, LIM_NOLIMIT3 = 2, LIM_VANALBADA2 = 3, LIM_MINMOD2 = 4,
    LIM_SUPERBEE2 = 5, LIM_MC2 = 6, LIM_VANLEER2 = 7, LIM_CADA3 = 8, LIM_SMOOTHVANALBADA = 9
};
enum BcId { BC_PERIODIC = 0, BC_RIEMANN = 1, BC_ENTROPY = 2, BC_FIXED = 3, BC_WALL = 4 };

// Mixed precision keeps the solution tile, face states and stored face fluxes in FP64 where the
// flux splits into an FP64 central part and a dissipation proportional to the jump (Euler with
// the central, Rusanov or Roe-family fluxes); SLAU and AUSM+up run their faces entirely in FP32.
constexpr bool mixed_wide(int law, int flux)
{
    return law == 1 && (flux == 0 || flux == 1 || flux == 2 || flux == 5 || flux == 6);
}

template <class R> struct NumLimits;
template <> struct NumLimits<double> {
    HC_HD static double min() { return DBL_MIN; }
    HC_HD static double max() { return DBL_MAX; }
};
template <> struct NumLimits<float> {
    HC_HD static float min() { return FLT_MIN; }
    HC_HD static float max() { return FLT_MAX; }
};

// Species<Euler,Real>, conservationLaws/euler/euler.h:48-60
template <class R> struct Species {
    R gamma, minf, lref, nu, pr, dt, Rg, gamma1;
};

// Surface metric as used by the flux functions: area and metric[0] (normal);
// metric[1] (tangent) is (-n1, n0) for every surface built by geom::surface
// (geometry/surface.ipp:12-29).
template <class R> struct FaceG {
    R area, n0, n1;
};

// State<Euler,2>, conservationLaws/euler/euler.h:62-83
template <class R> struct EState {
    R p, r, t, h, v2, a2, u, v;
};

// State<ScalarAdvection,2>, conservationLaws/scalarAdvection/scalarAdvection.h:48-59
template <class R> struct SState {
    R q, u, v;
};

template <int LAW> struct LawTraits;
template <> struct LawTraits<LAW_SCALAR> { static constexpr int nv = 3; };
template <> struct LawTraits<LAW_EULER> { static constexpr int nv = 4; };

template <class R> struct Flux4 {
    R f[4];
    R lambda;
};

HC_HD double hc_sqrt(double x) { return ::sqrt(x); }
HC_HD float hc_sqrt(float x) { return ::sqrtf(x); }
HC_HD double hc_fabs(double x) { return ::fabs(x); }
HC_HD float hc_fabs(float x) { return ::fabsf(x); }
HC_HD double hc_fmax(double a, double b) { return ::fmax(a, b); }
HC_HD float hc_fmax(float a, float b) { return ::fmaxf(a, b); }
HC_HD double hc_fmin(double a, double b) { return ::fmin(a, b); }
HC_HD float hc_fmin(float a, float b) { return ::fminf(a, b); }
HC_HD double hc_pow(double a, double b) { return ::pow(a, b); }
HC_HD float hc_pow(float a, float b) { return ::powf(a, b); }

// std::max / std::min / std::clamp semantics (first argument wins ties and NaN comparisons)
template <class R> HC_HD R std_max(R a, R b) { return (a < b) ? b : a; }
template <class R> HC_HD R std_clamp(R v, R lo, R hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// ---------------- set <-> state ----------------

// euler/transforms/conserved.ipp:31-76
template <class R> HC_HD EState<R> euler_from_conserved(const Species<R>& sp, const R* qc)
{
    EState<R> s;
    const R gam = sp.gamma;
    const R Rg = sp.Rg;
    const R gam1 = sp.gamma1;
    const R r = qc[2];
    const R re = qc[3];
    const R r1 = R(1.) / r;
    s.u = qc[0] * r1;
    s.v = qc[1] * r1;
    R v2 = R(0);
    v2 = v2 + s.u * s.u;
    v2 = v2 + s.v * s.v;
    s.v2 = v2;
    const R k = R(0.5) * v2;
    const R p = (gam - R(1.)) * (re - r * k);
    const R t = p / (Rg * r);
    const R a2 = gam * Rg * t;
    const R h = a2 * gam1 + k;
    s.p = p;
    s.r = r;
    s.t = t;
    s.h = h;
    s.a2 = a2;
    return s;
}

// euler/transforms/primitive.ipp:24-60
template <class R> HC_HD EState<R> euler_from_primitive(const Species<R>& sp, const R* qp)
{
    EState<R> s;
    const R gam = sp.gamma;
    const R Rg = sp.Rg;
    const R gam1 = sp.gamma1;
    R v2 = R(0);
    s.u = qp[0];
    v2 = v2 + qp[0] * qp[0];
    s.v = qp[1];
    v2 = v2 + qp[1] * qp[1];
    s.v2 = v2;
    const R k = R(0.5) * v2;
    const R r = qp[2];
    const R p = qp[3];
    const R t = p / (r * Rg);
    const R a2 = gam * Rg * t;
    const R h = a2 * gam1 + k;
    s.h = h;
    s.p = p;
    s.r = r;
    s.t = t;
    s.a2 = a2;
    return s;
}

// euler/transforms/conserved.ipp:2-29
template <class R> HC_HD void euler_to_conserved(const EState<R>& s, R* qc)
{
    const R re = s.r * s.h - s.p;
    qc[0] = s.r * s.u;
    qc[1] = s.r * s.v;
    qc[2] = s.r;
    qc[3] = re;
}

// euler/transforms/primitive.ipp:5-21
template <class R> HC_HD void euler_to_primitive(const EState<R>& s, R* qp)
{
    qp[0] = s.u;
    qp[1] = s.v;
    qp[2] = s.r;
    qp[3] = s.p;
}

template <class R, int BASIS> HC_HD EState<R> euler_state(const Species<R>& sp, const R* q)
{
    if constexpr (BASIS == BASIS_PRIMITIVE) return euler_from_primitive(sp, q);
    else return euler_from_conserved(sp, q);
}

// scalarAdvection/transforms/conserved.ipp:19-34
template <class R> HC_HD SState<R> scalar_state(const R* q)
{
    SState<R> s;
    s.u = q[0];
    s.v = q[1];
    s.q = q[2];
    return s;
}

// solution basis -> conserved (base/base.h:250-273 set2Set), in place into qc
template <class R, int LAW, int BASIS> HC_HD void to_conserved(const Species<R>& sp, const R* q, R* qc)
{
    if constexpr (LAW == LAW_EULER && BASIS == BASIS_PRIMITIVE) {
        const EState<R> s = euler_from_primitive(sp, q);
        euler_to_conserved(s, qc);
    } else {
        for (int v = 0; v < LawTraits<LAW>::nv; ++v) qc[v] = q[v];
    }
}

template <class R, int LAW, int BASIS> HC_HD void from_conserved(const Species<R>& sp, const R* qc, R* q)
{
    if constexpr (LAW == LAW_EULER && BASIS == BASIS_PRIMITIVE) {
        const EState<R> s = euler_from_conserved(sp, qc);
        euler_to_primitive(s, q);
    } else {
        for (int v = 0; v < LawTraits<LAW>::nv; ++v) q[v] = qc[v];
    }
}

// ---------------- delta transforms about a state ----------------

// euler/transforms/primitive.ipp:65-107
template <class R> HC_HD void dcons_to_dprim(const Species<R>& sp, const EState<R>& s, const R* dqc, R* dqp)
{
    const R dr = dqc[2];
    const R dre = dqc[3];
    const R rho1 = R(1.) / s.r;
    const R k = R(0.5) * s.v2;
    R c = k * dr + dre;
    {
        const R u = s.u;
        const R dru = dqc[0];
        dqp[0] = (dru - u * dr) * rho1;
        c = c - u * dru;
    }
    {
        const R u = s.v;
        const R dru = dqc[1];
        dqp[1] = (dru - u * dr) * rho1;
        c = c - u * dru;
    }
    const R dp = c * (sp.gamma - R(1.));
    dqp[2] = dr;
    dqp[3] = dp;
}

// euler/transforms/primitive.ipp:110-156
template <class R> HC_HD void dprim_to_dcons(const Species<R>& sp, const EState<R>& s, const R* dqp, R* dqc)
{
    const R dr = dqp[2];
    const R dp = dqp[3];
    const R r = s.r;
    R dk = R(0);
    {
        const R u = s.u;
        const R du = dqp[0];
        dqc[0] = r * du + u * dr;
        dk = dk + u * du;
    }
    {
        const R u = s.v;
        const R du = dqp[1];
        dqc[1] = r * du + u * dr;
        dk = dk + u * du;
    }
    const R k = R(0.5) * s.v2;
    const R dre = dp * sp.gamma1 + r * dk + k * dr;
    dqc[2] = dr;
    dqc[3] = dre;
}

// euler/transforms/characteristic.ipp:4-41
template <class R> HC_HD void dchar_to_dprim(const Species<R>& sp, const EState<R>& s, const R* dqw, R* dqp)
{
    const R r = s.r;
    const R a = hc_sqrt(s.a2);
    const R dp = R(0.5) * r * a * (dqw[2] - dqw[1]);
    const R dr = dp / (a * a) - r * dqw[0] / sp.gamma;
    dqp[0] = R(0.5) * (dqw[1] + dqw[2]);
    dqp[1] = dqw[3];
    dqp[2] = dr;
    dqp[3] = dp;
}

// euler/transforms/characteristic.ipp:43-79
template <class R> HC_HD void dprim_to_dchar(const Species<R>& sp, const EState<R>& s, const R* dqp, R* dqw)
{
    const R r1 = R(1.) / s.r;
    const R a1 = R(1.) / hc_sqrt(s.a2);
    const R du = dqp[0];
    const R dr = dqp[2];
    const R dp = dqp[3];
    dqw[0] = sp.gamma * (dp * a1 * a1 - dr) * r1;
    dqw[1] = du - dp * r1 * a1;
    dqw[2] = du + dp * r1 * a1;
    dqw[3] = dqp[1];
}

// ---------------- face-frame rotations (base/base.h:171-232) ----------------

// rotateToMetric: q1[i] = sum_j metric[i][j] q0[j] with metric = {n, t}, t = (-n1, n0)
template <class R> HC_HD void rotate_to_face(const FaceG<R>& f, const R* q0, R* q1)
{
    const R t0 = -f.n1, t1 = f.n0;
    R a = R(0);
    a = a + f.n0 * q0[0];
    a = a + f.n1 * q0[1];
    R b = R(0);
    b = b + t0 * q0[0];
    b = b + t1 * q0[1];
    q1[0] = a;
    q1[1] = b;
}

// rotateFromMetric: transposed metric
template <class R> HC_HD void rotate_from_face(const FaceG<R>& f, const R* q0, R* q1)
{
    const R t0 = -f.n1, t1 = f.n0;
    R a = R(0);
    a = a + f.n0 * q0[0];
    a = a + t0 * q0[1];
    R b = R(0);
    b = b + f.n1 * q0[0];
    b = b + t1 * q0[1];
    q1[0] = a;
    q1[1] = b;
}

// base/base.h:148-158 projectedVelocity
template <class R> HC_HD R proj_vel(R n0, R n1, R u, R v)
{
    R un = R(0);
    un = un + n0 * u;
    un = un + n1 * v;
    return un;
}

// ---------------- exact fluxes ----------------

// euler/fluxes/exactFlux.ipp:4-36
template <class R> HC_HD Flux4<R> euler_exact(const FaceG<R>& f, const EState<R>& s)
{
    Flux4<R> fr;
    const R a = hc_sqrt(s.a2);
    const R un = proj_vel(f.n0, f.n1, s.u, s.v);
    const R mn = s.r * un;
    fr.f[0] = mn * s.u + s.p * f.n0;
    fr.f[1] = mn * s.v + s.p * f.n1;
    fr.f[2] = mn;
    fr.f[3] = mn * s.h;
    fr.lambda = hc_fabs(un) + a;
    return fr;
}

// scalarAdvection/fluxes/exactFlux.ipp:8-21
template <class R> HC_HD Flux4<R> scalar_exact(const FaceG<R>& f, const SState<R>& s)
{
    Flux4<R> fr;
    const R un = proj_vel(f.n0, f.n1, s.u, s.v);
    fr.f[0] = R(0);
    fr.f[1] = R(0);
    fr.f[2] = un * s.q;
    fr.f[3] = R(0);
    fr.lambda = hc_fabs(un);
    return fr;
}

// euler/euler.h:161-166
template <class R> HC_HD R euler_specrad(R n0, R n1, const EState<R>& s)
{
    return hc_fabs(proj_vel(n0, n1, s.u, s.v)) + hc_sqrt(s.a2);
}

// ---------------- flux functors ----------------

// base/fluxes/centralFlux.ipp:2-14 (CentralFlux<Law>::flux, roeFlux.ipp:11-12)
template <class R> HC_HD Flux4<R> euler_central(const FaceG<R>& f, const EState<R>& sl, const EState<R>& sr)
{
    const Flux4<R> fl = euler_exact(f, sl);
    const Flux4<R> fr = euler_exact(f, sr);
    Flux4<R> o;
    for (int v = 0; v < 4; ++v) o.f[v] = ((fl.f[v] + fr.f[v]) * R(0.5)) * f.area;
    o.lambda = std_max(fl.lambda, fr.lambda) * f.area;
    return o;
}

template <class R> HC_HD Flux4<R> scalar_central(const FaceG<R>& f, const SState<R>& sl, const SState<R>& sr)
{
    const Flux4<R> fl = scalar_exact(f, sl);
    const Flux4<R> fr = scalar_exact(f, sr);
    Flux4<R> o;
    for (int v = 0; v < 3; ++v) o.f[v] = ((fl.f[v] + fr.f[v]) * R(0.5)) * f.area;
    o.f[3] = R(0);
    o.lambda = std_max(fl.lambda, fr.lambda) * f.area;
    return o;
}

// base/fluxes/rusanovFlux.ipp:2-50
template <class R> HC_HD Flux4<R> euler_rusanov(const FaceG<R>& f, const EState<R>& sl, const EState<R>& sr)
{
    Flux4<R> o = euler_central(f, sl, sr);
    const R lambda = std_max(euler_specrad(f.n0, f.n1, sl), euler_specrad(f.n0, f.n1, sr));
    R cr[4], cl[4];
    euler_to_conserved(sr, cr);
    euler_to_conserved(sl, cl);
    const R la = lambda * f.area;
    const R c = R(-0.5) * la;
    for (int v = 0; v < 4; ++v) o.f[v] = o.f[v] + (cr[v] - cl[v]) * c;
    return o;
}

template <class R> HC_HD Flux4<R> scalar_rusanov(const FaceG<R>& f, const SState<R>& sl, const SState<R>& sr)
{
    Flux4<R> o = scalar_central(f, sl, sr);
    const R lambda = std_max(hc_fabs(proj_vel(f.n0, f.n1, sl.u, sl.v)), hc_fabs(proj_vel(f.n0, f.n1, sr.u, sr.v)));
    const R cr[3] = {sr.u, sr.v, sr.q};
    const R cl[3] = {sl.u, sl.v, sl.q};
    const R la = lambda * f.area;
    const R c = R(-0.5) * la;
    for (int v = 0; v < 3; ++v) o.f[v] = o.f[v] + (cr[v] - cl[v]) * c;
    return o;
}

// euler/fluxes/roe.ipp:2-48
template <class R> HC_HD EState<R> roe_average(const Species<R>& sp, const EState<R>& sl, const EState<R>& sr)
{
    EState<R> a;
    const R cr = hc_sqrt(sr.r / sl.r);
    const R wl = R(1.) / (R(1.) + cr);
    const R wr = R(1.) - wl;
    const R ra = cr * sl.r;
    R ka = R(0);
    const R ua = wl * sl.u + wr * sr.u;
    ka = ka + ua * ua;
    const R va = wl * sl.v + wr * sr.v;
    ka = ka + va * va;
    ka = ka * R(0.5);
    a.u = ua;
    a.v = va;
    const R ha = wl * sl.h + wr * sr.h;
    const R a2a = (sp.gamma - R(1.)) * (ha - ka);
    const R ta = a2a / (sp.Rg * sp.gamma);
    const R pa = ra * sp.Rg * ta;
    a.p = pa;
    a.r = ra;
    a.t = ta;
    a.h = ha;
    a.v2 = R(2.) * ka;
    a.a2 = a2a;
    return a;
}

// euler/euler.h:168-181 wavespeeds {u, u-a, u+a, u}
template <class R> HC_HD void wavespeeds(const FaceG<R>& f, const EState<R>& s, R* w)
{
    const R u = proj_vel(f.n0, f.n1, s.u, s.v);
    const R a = hc_sqrt(s.a2);
    w[0] = u;
    w[1] = u - a;
    w[2] = u + a;
    w[3] = u;
}

// euler/fluxes/roe.ipp:50-78 (Harten entropy fix, eps = 0.05)
template <class R> HC_HD void entropyfix(const R* la, const R* ll, const R* lr, R* s)
{
    const R ua = hc_fabs(la[0]);
    s[0] = ua;
    s[3] = ua;
    const R eps = R(0.05);
    for (int i = 1; i < 3; i++) {
        const R le = eps * hc_fmax(R(0.), hc_fmax(la[i] - ll[i], lr[i] - la[i]));
        const R laa = hc_fabs(la[i]);
        s[i] = (laa > le) ? laa : R(0.5) * (le + laa * laa / le);
        // FP32 (the mixed path's dissipation): a sonic Roe state (laa = 0) inside a compression
        // (le = 0) is frequent enough at 24 bits to hit 0/0; the formula's limit there is laa
        if constexpr (sizeof(R) == sizeof(float)) {
            if (!(le > R(0))) s[i] = laa;
        }
    }
}

// base/fluxes/roeFlux.ipp:2-58 (RoeDissipation): fd = R |Lambda| L (dq) in the global frame for the
// conserved jump dq, returns the Roe-state spectral radius times the area.  FACEFRAME = true is the
// added variant that rotates the Roe-average velocity into the face frame like the jump; the
// reference keeps it in the global frame (roeFlux.ipp:40-54).
template <class R, bool FACEFRAME = false>
HC_HD R roe_dissipation(const Species<R>& sp, const FaceG<R>& f, const EState<R>& sl, const EState<R>& sr, const R* dq,
                        R* fd)
{
    const EState<R> ravg = roe_average(sp, sl, sr);
    EState<R> rv = ravg;
    if constexpr (FACEFRAME) {
        const R uv[2] = {ravg.u, ravg.v};
        R a[2];
        rotate_to_face(f, uv, a);
        rv.u = a[0];
        rv.v = a[1];
    }
    R wa[4], wl[4], wr[4], lam[4];
    wavespeeds(f, ravg, wa);
    wavespeeds(f, sl, wl);
    wavespeeds(f, sr, wr);
    entropyfix(wa, wl, wr, lam);
    R dqc[4];
    rotate_to_face(f, dq, dqc);
    dqc[2] = dq[2];
    dqc[3] = dq[3];
    R dqp[4], dqw[4];
    dcons_to_dprim(sp, rv, dqc, dqp);
    dprim_to_dchar(sp, rv, dqp, dqw);
    for (int v = 0; v < 4; ++v) dqw[v] = lam[v] * dqw[v];
    R dqp2[4], dqc2[4];
    dchar_to_dprim(sp, rv, dqw, dqp2);
    dprim_to_dcons(sp, rv, dqp2, dqc2);
    rotate_from_face(f, dqc2, fd);
    fd[2] = dqc2[2];
    fd[3] = dqc2[3];
    return euler_specrad(f.n0, f.n1, ravg) * f.area;
}

// base/fluxes/roeFlux.ipp:2-58 (RoeFlux = CentralFlux + RoeDissipation)
template <class R, bool FACEFRAME = false>
HC_HD Flux4<R> euler_roe(const Species<R>& sp, const FaceG<R>& f, const EState<R>& sl, const EState<R>& sr)
{
    const Flux4<R> c = euler_central(f, sl, sr);
    R cr[4], cl[4], dq[4], fd[4];
    euler_to_conserved(sr, cr);
    euler_to_conserved(sl, cl);
    for (int v = 0; v < 4; ++v) dq[v] = cr[v] - cl[v];
    const R dl = roe_dissipation<R, FACEFRAME>(sp, f, sl, sr, dq, fd);
    const R m = R(-0.5) * f.area;
    Flux4<R> o;
    for (int v = 0; v < 4; ++v) o.f[v] = c.f[v] + fd[v] * m;
    o.lambda = hc_fmax(c.lambda, dl);
    return o;
}

// euler/fluxes/roeUnprecWS.ipp:24-104 (RoeUnprecWS, the "Weiss & Smith / Li & Gu form" of
// euler.h:392-405), dissipation part: fd = 0.5 A (|u_n| dq + dU (rho u, rho v, rho, rho H) +
// dP (n_x, n_y, 0, u_n)) of the Roe state, dU = nu_p dp + nu_u du, dP = mu_p dp + mu_u du, with
// the source's fixed coefficients and exponents A11 = A21 = A22 = 1, A12 = 0 on the species'
// free-stream Mach number (pow(minf, 1) = minf, pow(minf, 0) = 1, both exact); returns the Roe
// state's spectral radius times the area.  dp, du and the conserved jump dq come in from the caller.
template <class R>
HC_HD R roe_unprec_ws_dissipation(const Species<R>& sp, const FaceG<R>& f, const EState<R>& ravg, R dp, R du,
                                  const R* dq, R* fd)
{
    const R una = proj_vel(f.n0, f.n1, ravg.u, ravg.v);
    const R ra = ravg.r;
    const R aa = hc_sqrt(ravg.a2);
    const R minf = sp.minf;
    const R uinf = minf * aa;
    const R nu_p = R(1.00) * (R(1.) / (ra * uinf * uinf)) * minf;
    const R nu_u = R(0.05);
    const R mu_p = R(0.02) * minf;
    const R mu_u = R(0.10) * (ra * uinf) * minf;
    const R dU = nu_p * dp + nu_u * du;
    const R dP = mu_p * dp + mu_u * du;
    const R ruvh[4] = {ra * ravg.u, ra * ravg.v, ra, ra * ravg.h};
    const R onu[4] = {f.n0, f.n1, R(0), una};
    const R aun = hc_fabs(una);
    for (int v = 0; v < 4; ++v) fd[v] = (R(0.5) * ((aun * dq[v] + dU * ruvh[v]) + dP * onu[v])) * f.area;
    return euler_specrad(f.n0, f.n1, ravg) * f.area;
}

// euler/fluxes/roeUnprecWS.ipp:1-109: central flux minus the dissipation above,
// lambda = max(central, Roe-state spectral radius * area)
template <class R>
HC_HD Flux4<R> euler_roe_unprec_ws(const Species<R>& sp, const FaceG<R>& f, const EState<R>& sl, const EState<R>& sr)
{
    const EState<R> ravg = roe_average(sp, sl, sr);
    const Flux4<R> c = euler_central(f, sl, sr);
    const R dp = sr.p - sl.p;
    const R du = proj_vel(f.n0, f.n1, sr.u, sr.v) - proj_vel(f.n0, f.n1, sl.u, sl.v);
    R cr[4], cl[4], dq[4], fd[4];
    euler_to_conserved(sr, cr);
    euler_to_conserved(sl, cl);
    for (int v = 0; v < 4; ++v) dq[v] = cr[v] - cl[v];
    const R dl = roe_unprec_ws_dissipation(sp, f, ravg, dp, du, dq, fd);
    Flux4<R> o;
    for (int v = 0; v < 4; ++v) o.f[v] = c.f[v] - fd[v];
    o.lambda = hc_fmax(c.lambda, dl);
    return o;
}

template <class R> HC_HD EState<float> state_f32(const EState<R>& s)
{
    return EState<float>{float(s.p), float(s.r), float(s.t), float(s.h), float(s.v2), float(s.a2), float(s.u), float(s.v)};
}

// Mixed-precision Roe: the central flux and the conserved jump in FP64 (O(1) quantities), the
// dissipation R |Lambda| L (dq) in FP32 (proportional to the jump, so FP32 round-off stays
// relative to the jump and not to the flux)
template <bool FACEFRAME>
HC_HD Flux4<double> euler_roe_mixed(const Species<float>& spf, const FaceG<double>& f, const EState<double>& sl,
                                    const EState<double>& sr)
{
    const Flux4<double> c = euler_central(f, sl, sr);
    double cr[4], cl[4];
    euler_to_conserved(sr, cr);
    euler_to_conserved(sl, cl);
    float dq[4], fd[4];
    for (int v = 0; v < 4; ++v) dq[v] = float(cr[v] - cl[v]);
    const FaceG<float> ff{float(f.area), float(f.n0), float(f.n1)};
    const float dl = roe_dissipation<float, FACEFRAME>(spf, ff, state_f32(sl), state_f32(sr), dq, fd);
    const double m = -0.5 * f.area;
    Flux4<double> o;
    for (int v = 0; v < 4; ++v) o.f[v] = c.f[v] + double(fd[v]) * m;
    o.lambda = hc_fmax(c.lambda, double(dl));
    return o;
}

// Mixed-precision RoeUnprecWS: central flux and the jumps (dq, dp, du) in FP64, the dissipation
// in FP32 on the Roe state of the rounded face states
HC_HD Flux4<double> euler_roe_unprec_ws_mixed(const Species<float>& spf, const FaceG<double>& f,
                                              const EState<double>& sl, const EState<double>& sr)
{
    const Flux4<double> c = euler_central(f, sl, sr);
    double cr[4], cl[4];
    euler_to_conserved(sr, cr);
    euler_to_conserved(sl, cl);
    float dq[4], fd[4];
    for (int v = 0; v < 4; ++v) dq[v] = float(cr[v] - cl[v]);
    const float dp = float(sr.p - sl.p);
    const float du = float(proj_vel(f.n0, f.n1, sr.u, sr.v) - proj_vel(f.n0, f.n1, sl.u, sl.v));
    const FaceG<float> ff{float(f.area), float(f.n0), float(f.n1)};
    const EState<float> ravg = roe_average(spf, state_f32(sl), state_f32(sr));
    const float dl = roe_unprec_ws_dissipation(spf, ff, ravg, dp, du, dq, fd);
    Flux4<double> o;
    for (int v = 0; v < 4; ++v) o.f[v] = c.f[v] - double(fd[v]);
    o.lambda = hc_fmax(c.lambda, double(dl));
    return o;
}

// euler/fluxes/machSplittings.ipp:8-43
template <class R> HC_HD R ms_m1(int sign, R M) { return R(0.5) * (M + R(sign) * hc_fabs(M)); }
template <class R> HC_HD R ms_m2(int sign, R M) { return R(sign) * R(0.25) * (M + R(sign)) * (M + R(sign)); }
template <class R> HC_HD R ms_m4(int sign, R M, R beta)
{
    return ms_m2(sign, M) * (R(1.) - R(sign) * R(16.) * beta * ms_m2(-sign, M));
}
template <class R> HC_HD R ms_p1(int sign, R M)
{
    const R eps = NumLimits<R>::max();
    return (ms_m1(sign, M) + eps) / (hc_fabs(M) + eps);
}
template <class R> HC_HD R ms_p3(int sign, R M) { return ms_m2(sign, M) * (R(sign) * R(2.) - M); }
template <class R> HC_HD R ms_p5(int sign, R M, R alpha)
{
    return ms_p3(sign, M) - R(sign) * R(16.) * alpha * M * ms_m2(sign, M) * ms_m2(-sign, M);
}

// euler/fluxes/slau.ipp:4-161 (SLAU mass flux, SLAU2 pressure flux)
template <class R> HC_HD Flux4<R> euler_slau(const Species<R>& sp, const FaceG<R>& f, const EState<R>& sl, const EState<R>& sr)
{
    const R ua = hc_sqrt(R(0.5) * (sl.v2 + sr.v2));
    const R aa = R(0.5) * (hc_sqrt(sl.a2) + hc_sqrt(sr.a2));
    const R aa1 = R(1.0) / aa;
    const R ma = ua * aa1;
    const R ma_hat = hc_fmin(R(1.0), ma);
    const R unl = proj_vel(f.n0, f.n1, sl.u, sl.v);
    const R unr = proj_vel(f.n0, f.n1, sr.u, sr.v);
    const R ml = unl * aa1;
    const R mr = unr * aa1;
    const R una = (sl.r * hc_fabs(unl) + sr.r * hc_fabs(unr)) / (sl.r + sr.r);
    const R chi = (R(1.) - ma_hat) * (R(1.) - ma_hat);
    const R g = -hc_fmax(hc_fmin(ml, R(0.)), R(-1.)) * hc_fmin(hc_fmax(mr, R(0.)), R(1.));
    const R dr = sr.r - sl.r;
    const R dp = sr.p - sl.p;
    const R mdot_c = sl.r * unl + sr.r * unr;
    const R mdot_dr = una * dr;
    const R mdot_dp = chi * dp * aa1;
    const R mdot = R(0.5) * ((mdot_c - mdot_dr) * (R(1.) - g) - mdot_dp) * f.area;
    const R mlsup = hc_fabs(ml) > R(1) ? R(1) : R(0);
    const R mlsub = R(1) - mlsup;
    const R mrsup = hc_fabs(mr) > R(1) ? R(1) : R(0);
    const R mrsub = R(1) - mrsup;
    const R betal = mlsub * ms_p3(1, ml) + mlsup * ms_p1(1, ml);
    const R betar = mrsub * ms_p3(-1, mr) + mrsup * ms_p1(-1, mr);
    const R p_c = (sl.p + sr.p);
    const R p_dp = (betal - betar) * dp;
    const R p_du = -ma * (betal + betar - R(1.)) * (p_c);
    const R pa = R(0.5) * (p_c - p_dp - p_du) * f.area;
    const R m1l = (ua > R(0)) ? R(1) : R(0);
    const R m1r = (ua > R(0)) ? R(0) : R(1);
    Flux4<R> o;
    o.f[0] = mdot * (m1l * sl.u + m1r * sr.u) + pa * f.n0;
    o.f[1] = mdot * (m1l * sl.v + m1r * sr.v) + pa * f.n1;
    o.f[2] = mdot;
    o.f[3] = mdot * (m1l * sl.h + m1r * sr.h);
    const R lmax = hc_sqrt(hc_fmax(sl.a2, sr.a2)) + hc_fmax(hc_fabs(unl), hc_fabs(unr));
    o.lambda = lmax * f.area;
    return o;
}

// euler/fluxes/ausmPlusUP.ipp:6-171
template <class R> HC_HD Flux4<R> euler_ausmpup(const Species<R>& sp, int vscale, int pscale, const FaceG<R>& f,
                                                const EState<R>& sl, const EState<R>& sr)
{
    const R alpha0 = R(0.1875);
    const R beta = R(0.125);
    const R Ku = R(0.75) * R(2.);
    const R Kp = R(0.25);
    const R sigma = R(1.0);
    const R ascoeff = R(2.) * (sp.gamma - R(1.)) / (sp.gamma + R(1.));
    const R unl = proj_vel(f.n0, f.n1, sl.u, sl.v);
    const R unr = proj_vel(f.n0, f.n1, sr.u, sr.v);
    const R als2 = ascoeff * sl.h;
    const R ars2 = ascoeff * sr.h;
    const R als = als2 / hc_fmax(unl, hc_sqrt(als2));
    const R ars = ars2 / hc_fmax(-unr, hc_sqrt(ars2));
    const R ra = R(0.5) * (sl.r + sr.r);
    const R as = hc_fmin(ars, als);
    const R ml = unl / as;
    const R mr = unr / as;
    const R minf = sp.minf;
    const R m0 = hc_sqrt(R(0.5) * (ml * ml + mr * mr));
    const R mu = sp.lref / (as * sp.dt);
    const R m0_c = std_clamp(m0, minf, R(1.));
    const R m0_a = std_clamp(hc_fmax(m0, mu), minf, R(1.));
    const R m0_p = pscale == SCALE_CONVECTIVE ? m0_c : pscale == SCALE_ACOUSTIC ? R(1.0) : m0_a;
    const R m0_u = vscale == SCALE_CONVECTIVE ? m0_c : vscale == SCALE_ACOUSTIC ? R(1.0) : m0_a;
    const R fa_p = m0_p * (R(2.) - m0_p);
    const R fa_u = m0_u * (R(2.) - m0_u);
    const R mlsup = hc_fabs(ml) > R(1) ? R(1) : R(0);
    const R mlsub = R(1) - mlsup;
    const R mrsup = hc_fabs(mr) > R(1) ? R(1) : R(0);
    const R mrsub = R(1) - mrsup;
    const R mlp = mlsub * ms_m4(1, ml, beta) + mlsup * ms_m1(1, ml);
    const R mrm = mrsub * ms_m4(-1, mr, beta) + mrsup * ms_m1(-1, mr);
    const R delp = sr.p - sl.p;
    const R mp = Kp * hc_fmax(R(1.) - sigma * m0 * m0, R(0.)) * delp / (fa_p * ra * as * as);
    const R ua = as * (mlp + mrm - mp) * f.area;
    const R alpha = alpha0 * (R(5.) * fa_u * fa_u - R(4.));
    const R plp = mlsub * ms_p5(1, ml, alpha) + mlsup * ms_p1(1, ml);
    const R prm = mrsub * ms_p5(-1, mr, alpha) + mrsup * ms_p1(-1, mr);
    const R delu = unr - unl;
    const R pu = Ku * plp * prm * ra * (fa_u * as) * delu;
    const R pa = (plp * sl.p + prm * sr.p - pu) * f.area;
    const R m1l = (ua > R(0)) ? R(1) : R(0);
    const R m1r = (ua > R(0)) ? R(0) : R(1);
    const R mdot = ua * (m1l * sl.r + m1r * sr.r);
    Flux4<R> o;
    o.f[0] = mdot * (m1l * sl.u + m1r * sr.u) + pa * f.n0;
    o.f[1] = mdot * (m1l * sl.v + m1r * sr.v) + pa * f.n1;
    o.f[2] = mdot;
    o.f[3] = mdot * (m1l * sl.h + m1r * sr.h);
    const R lmd = R(0.5) * (m0_p + R(1)) / fa_p;
    const R amax = hc_sqrt(hc_fmax(as * as, hc_fmax(sl.a2, sr.a2)));
    const R lmax = lmd * amax + hc_fmax(hc_fabs(ua), hc_fmax(hc_fabs(unl), hc_fabs(unr)));
    o.lambda = lmax * f.area;
    return o;
}

// FluxInterface (base/base.h:370-397): solution sets -> states -> Flux::flux
template <class R, int LAW, int BASIS, int FLUX>
HC_HD Flux4<R> face_flux(const Species<R>& sp, int vscale, int pscale, const FaceG<R>& f, const R* ql, const R* qr)
{
    if constexpr (LAW == LAW_SCALAR) {
        const SState<R> sl = scalar_state(ql);
        const SState<R> sr = scalar_state(qr);
        if constexpr (FLUX == FLUX_CENTRAL) return scalar_central(f, sl, sr);
        else return scalar_rusanov(f, sl, sr);
    } else {
        const EState<R> sl = euler_state<R, BASIS>(sp, ql);
        const EState<R> sr = euler_state<R, BASIS>(sp, qr);
        if constexpr (FLUX == FLUX_CENTRAL) return euler_central(f, sl, sr);
        else if constexpr (FLUX == FLUX_RUSANOV) return euler_rusanov(f, sl, sr);
        else if constexpr (FLUX == FLUX_ROE) return euler_roe(sp, f, sl, sr);
        else if constexpr (FLUX == FLUX_SLAU) return euler_slau(sp, f, sl, sr);
        else if constexpr (FLUX == FLUX_ROE_FACEFRAME) return euler_roe<R, true>(sp, f, sl, sr);
        else if constexpr (FLUX == FLUX_ROE_UNPREC_WS) return euler_roe_unprec_ws(sp, f, sl, sr);
        else return euler_ausmpup(sp, vscale, pscale, f, sl, sr);
    }
}

// Mixed-precision face flux on FP64 face states: the Roe family (Roe, face-frame Roe,
// RoeUnprecWS) splits into an FP64 central flux and an FP32 dissipation; Rusanov and central are cheap and stay FP64; SLAU and AUSM+up, which
// do not split that way, run in FP32 on the rounded face states.
template <int LAW, int BASIS, int FLUX>
HC_HD Flux4<double> face_flux_mixed(const Species<double>& sp, const Species<float>& spf, int vscale, int pscale,
                                    const FaceG<double>& f, const double* ql, const double* qr)
{
    if constexpr (LAW == LAW_EULER && (FLUX == FLUX_ROE || FLUX == FLUX_ROE_FACEFRAME)) {
        const EState<double> sl = euler_state<double, BASIS>(sp, ql);
        const EState<double> sr = euler_state<double, BASIS>(sp, qr);
        return euler_roe_mixed<FLUX == FLUX_ROE_FACEFRAME>(spf, f, sl, sr);
    } else if constexpr (LAW == LAW_EULER && FLUX == FLUX_ROE_UNPREC_WS) {
        const EState<double> sl = euler_state<double, BASIS>(sp, ql);
        const EState<double> sr = euler_state<double, BASIS>(sp, qr);
        return euler_roe_unprec_ws_mixed(spf, f, sl, sr);
    } else if constexpr (FLUX == FLUX_RUSANOV || FLUX == FLUX_CENTRAL) {
        return face_flux<double, LAW, BASIS, FLUX>(sp, vscale, pscale, f, ql, qr);
    } else {
        constexpr int NV = LawTraits<LAW>::nv;
        float a[NV], b[NV];
        for (int v = 0; v < NV; ++v) {
            a[v] = float(ql[v]);
            b[v] = float(qr[v]);
        }
        const FaceG<float> ff{float(f.area), float(f.n0), float(f.n1)};
        const Flux4<float> F = face_flux<float, LAW, BASIS, FLUX>(spf, vscale, pscale, ff, a, b);
        Flux4<double> o;
        for (int v = 0; v < 4; ++v) o.f[v] = double(F.f[v]);
        o.lambda = double(F.lambda);
        return o;
    }
}

// ---------------- limiters (limiters/limiter.ipp) ----------------

// limiter L of the pair (a, b): a the centred slope, b the upwind-biased one
template <int L, class R> HC_HD R limit_fn(R a, R b)
{
    if constexpr (L == LIM_NOLIMIT1) {   // limiter.ipp:9-13
        return R(0);
    } else if constexpr (L == LIM_NOLIMIT2) {   // limiter.ipp:15-19
        return R(0.5) * (a + b);
    } else if constexpr (L == LIM_NOLIMIT3) {   // limiter.ipp:21-25
        return R(0.33333333333333333) * (R(2.) * a + b);
    } else if constexpr (L == LIM_VANALBADA2) {   // limiter.ipp:27-38
        const R eps = NumLimits<R>::min();
        return a * b < R(0) ? R(0) : (b * (a * a) + a * (b * b)) / (a * a + b * b + eps);
    } else if constexpr (L == LIM_MINMOD2) {   // limiter.ipp:40-48
        return a > R(0) ? hc_fmin(a, hc_fmax(b, R(0.))) : hc_fmax(a, hc_fmin(b, R(0.)));
    } else if constexpr (L == LIM_SUPERBEE2) {   // limiter.ipp:50-63
        return a > R(0) ? hc_fmax(R(0.), hc_fmax(hc_fmin(R(2) * a, b), hc_fmin(R(2) * b, a)))
                        : hc_fmin(R(0.), hc_fmin(hc_fmax(R(2) * a, b), hc_fmax(R(2) * b, a)));
    } else if constexpr (L == LIM_MC2) {   // limiter.ipp:65-80
        return a > R(0) ? hc_fmax(R(0.), hc_fmin(R(0.5) * (a + b), hc_fmin(R(2) * b, R(2) * a)))
                        : hc_fmin(R(0.), hc_fmax(R(0.5) * (a + b), hc_fmax(R(2) * b, R(2) * a)));
    } else if constexpr (L == LIM_VANLEER2) {   // limiter.ipp:82-90
        const R eps = NumLimits<R>::min();
        return (b + hc_fabs(b)) / (a + hc_fabs(b) + eps);
    } else if constexpr (L == LIM_CADA3) {   // limiter.ipp:92-161
        const R t = R(0.0);
        const R gam = R(1.5);
        const R c1 = R(-2. / 7.);
        const R c2 = R(0.4);
        const R c3 = R(3) * gam - R(2);
        R u, v, w;
        if (a > R(0)) { u = a; v = b; }
        else { u = -a; v = -b; }
        const R p = (v + R(2.) * u) / R(3.);
        const R q = hc_fabs(a) + hc_fabs(b);
        if (q > t) {
            if (v < R(-2) * u) w = R(0);
            else if (v < c1 * u) w = (v + R(2.) * u) / R(3.);
            else if (v < R(0)) w = R(-2.) * v;
            else if (v < c2 * u) w = R(2.) * v;
            else if (v < c3 * u) w = (v + R(2.) * u) / R(3.);
            else w = gam * u;
        } else {
            w = p;
        }
        if (a <= R(0)) w = -w;
        return w;
    } else {   // smooth van Albada: VanAlbada2 without the a*b<0 switch
        const R eps = NumLimits<R>::min();
        return (b * (a * a) + a * (b * b)) / (a * a + b * b + eps);
    }
}

#define HC_FOR_EACH_LIMITER(X)                                                                            \
    X(LIM_NOLIMIT1) X(LIM_NOLIMIT2) X(LIM_NOLIMIT3) X(LIM_VANALBADA2) X(LIM_MINMOD2) X(LIM_SUPERBEE2) X(LIM_MC2) \
    X(LIM_VANLEER2) X(LIM_CADA3)

// limiters/limiter.ipp:10-161 by run-time id (an unknown id is the smooth van Albada limiter)
template <class R> HC_HD R limit(int lim, R a, R b)
{
    switch (lim) {
#define HC_LIM_CASE(L) \
    case L: return limit_fn<L>(a, b);
        HC_FOR_EACH_LIMITER(HC_LIM_CASE)
#undef HC_LIM_CASE
    default: return limit_fn<LIM_SMOOTHVANALBADA>(a, b);
    }
}

// the limiter of every pair (a[v], bl[v]) and (a[v], br[v]), v < N: one branch on the id for the
// face, then straight-line code over the pairs (the same operations as limit, pair by pair)
template <class R, int N> HC_HD void limit_pairs(int lim, const R* a, const R* bl, const R* br, R* sl, R* sr)
{
    switch (lim) {
#define HC_LIM_CASE(L)                           \
    case L:                                      \
        for (int v = 0; v < N; ++v) {            \
            sl[v] = limit_fn<L>(a[v], bl[v]);    \
            sr[v] = limit_fn<L>(a[v], br[v]);    \
        }                                        \
        return;
        HC_FOR_EACH_LIMITER(HC_LIM_CASE)
#undef HC_LIM_CASE
    default:
        for (int v = 0; v < N; ++v) {
            sl[v] = limit_fn<LIM_SMOOTHVANALBADA>(a[v], bl[v]);
            sr[v] = limit_fn<LIM_SMOOTHVANALBADA>(a[v], br[v]);
        }
    }
}

// ---------------- boundary point functions ----------------

// Inviscid wall flux with q_wall = q_interior, euler/boundary/inviscidWall.ipp:33-72.
// f is the surface with its normal into the domain.
template <class R, int BASIS> HC_HD Flux4<R> euler_wall_flux(const Species<R>& sp, const FaceG<R>& f, const R* qint)
{
    const EState<R> s = euler_state<R, BASIS>(sp, qint);
    const R pw = s.p;
    Flux4<R> o;
    o.f[0] = f.area * pw * f.n0;
    o.f[1] = f.area * pw * f.n1;
    o.f[2] = R(0.);
    o.f[3] = R(0.);
    o.lambda = f.area * euler_specrad(f.n0, f.n1, s);
    return o;
}

template <class R, int BASIS> HC_HD void rotated_state_in(const FaceG<R>& f, const R* q, R* out)
{
    rotate_to_face(f, q, out);
    out[2] = q[2];
    out[3] = q[3];
}

template <class R, int BASIS> HC_HD void prim_to_solution(const Species<R>& sp, const R* qb, R* qs)
{
    if constexpr (BASIS == BASIS_PRIMITIVE) {
        for (int v = 0; v < 4; ++v) qs[v] = qb[v];
    } else {
        const EState<R> s = euler_from_primitive(sp, qb);
        euler_to_conserved(s, qs);
    }
}

// euler/boundary/riemann.ipp:18-113 (f: normal into the domain).  Returns 0 when
// the face normal does not point into the domain (the reference asserts).
template <class R, int BASIS> HC_HD int euler_riemann(const Species<R>& sp, const FaceG<R>& f, R fcx, R fcy, R ccx, R ccy,
                                                      const R* qint, const R* qfar, R* out)
{
    const R gam = sp.gamma;
    const R gam1 = sp.gamma1;
    const R d0 = ccx - fcx, d1 = ccy - fcy;
    R dd = R(0);
    dd = dd + d0 * f.n0;
    dd = dd + d1 * f.n1;
    const int ok = dd > R(0);
    R ri[4], rf[4];
    rotated_state_in<R, BASIS>(f, qint, ri);
    rotated_state_in<R, BASIS>(f, qfar, rf);
    const EState<R> si = euler_state<R, BASIS>(sp, ri);
    const EState<R> sf = euler_state<R, BASIS>(sp, rf);
    const EState<R> sa = roe_average(sp, si, sf);
    if ((sa.u / hc_sqrt(sa.v2)) < R(0.001)) {
        for (int v = 0; v < 4; ++v) out[v] = qint[v];
        return ok;
    }
    const R ma = sa.u / hc_sqrt(sa.a2);
    const bool inflow = ma > R(0);
    const R* quw = inflow ? qfar : qint;
    if (hc_fabs(ma) > R(1.)) {
        for (int v = 0; v < 4; ++v) out[v] = quw[v];
        return ok;
    }
    const R ai = hc_sqrt(si.a2);
    const R af = hc_sqrt(sf.a2);
    const R uni = si.u;
    const R unf = sf.u;
    const R rim = uni - R(2.) * ai * gam1;
    const R rip = unf + R(2.) * af * gam1;
    const R ub = R(0.5) * (rim + rip);
    const R ab = R(0.25) * (gam - R(1.)) * (rip - rim);
    R ru[4];
    rotated_state_in<R, BASIS>(f, quw, ru);
    const EState<R> su = euler_state<R, BASIS>(sp, ru);
    const R sb = su.a2 / (gam * hc_pow(su.r, gam - R(1.)));
    const R vb = su.v;
    const R rb = hc_pow(ab * ab / (gam * sb), gam1);
    const R pb = rb * ab * ab / gam;
    const R qb[4] = {ub, vb, rb, pb};
    R qs[4];
    prim_to_solution<R, BASIS>(sp, qb, qs);
    rotate_from_face(f, qs, out);
    out[2] = qs[2];
    out[3] = qs[3];
    return ok;
}

// euler/boundary/entropy.ipp:18-88
template <class R, int BASIS> HC_HD int euler_entropy(const Species<R>& sp, const FaceG<R>& f, R fcx, R fcy, R ccx, R ccy,
                                                      const R* qint, const R* qfar, R* out)
{
    const R gam = sp.gamma;
    const R d0 = ccx - fcx, d1 = ccy - fcy;
    R dd = R(0);
    dd = dd + d0 * f.n0;
    dd = dd + d1 * f.n1;
    const int ok = dd > R(0);
    R ri[4], rf[4];
    rotated_state_in<R, BASIS>(f, qint, ri);
    rotated_state_in<R, BASIS>(f, qfar, rf);
    const EState<R> si = euler_state<R, BASIS>(sp, ri);
    const EState<R> sf = euler_state<R, BASIS>(sp, rf);
    const EState<R> sa = roe_average(sp, si, sf);
    if ((sa.u / hc_sqrt(sa.v2)) < R(0.0001)) {
        for (int v = 0; v < 4; ++v) out[v] = qint[v];
        return ok;
    }
    const R ma = sa.u / hc_sqrt(sa.a2);
    const bool inflow = ma > R(0);
    const R* quw = inflow ? qfar : qint;
    const R* qdw = inflow ? qint : qfar;
    if (hc_fabs(ma) > R(1.)) {
        for (int v = 0; v < 4; ++v) out[v] = quw[v];
        return ok;
    }
    R ru[4], rd[4];
    rotated_state_in<R, BASIS>(f, quw, ru);
    rotated_state_in<R, BASIS>(f, qdw, rd);
    const EState<R> su = euler_state<R, BASIS>(sp, ru);
    const EState<R> sd = euler_state<R, BASIS>(sp, rd);
    const R pb = sd.p;
    const R sb = su.a2 / (gam * hc_pow(su.r, gam - R(1.)));
    const R rb = hc_pow(pb / sb, R(1.) / gam);
    const R qb[4] = {su.u, su.v, rb, pb};
    R qs[4];
    prim_to_solution<R, BASIS>(sp, qb, qs);
    rotate_from_face(f, qs, out);
    out[2] = qs[2];
    out[3] = qs[3];
    return ok;
}

// scalarAdvection/boundary/riemann.ipp:15-40
template <class R> HC_HD void scalar_riemann(const FaceG<R>& f, const R* qint, const R* qref, R* out)
{
    R avg[3], rot[3];
    for (int v = 0; v < 3; ++v) avg[v] = qint[v] + (qref[v] - qint[v]) * R(0.5);
    rotate_to_face(f, avg, rot);
    rot[2] = avg[2];
    const bool inflow = proj_vel(f.n0, f.n1, rot[0], rot[1]) > R(0);
    for (int v = 0; v < 3; ++v) out[v] = inflow ? qref[v] : qint[v];
}

// Ghost value of a non-periodic boundary (spatial/boundary/boundaryUpdate.h:137-266,
// spatial/boundary/boundaryCondition.h:63-80 for Fixed).  f has its normal into the domain.
template <class R, int LAW, int BASIS>
HC_HD int ghost_value(const Species<R>& sp, int bctype, const FaceG<R>& f, R fcx, R fcy, R ccx, R ccy,
                      const R* qint, const R* qref, R* out)
{
    constexpr int NV = LawTraits<LAW>::nv;
    if (bctype == BC_RIEMANN) {
        if constexpr (LAW == LAW_SCALAR) {
            scalar_riemann(f, qint, qref, out);
            return 1;
        } else {
            return euler_riemann<R, BASIS>(sp, f, fcx, fcy, ccx, ccy, qint, qref, out);
        }
    }
    if constexpr (LAW == LAW_EULER) {
        if (bctype == BC_ENTROPY) return euler_entropy<R, BASIS>(sp, f, fcx, fcy, ccx, ccy, qint, qref, out);
    }
    for (int v = 0; v < NV; ++v) out[v] = qref[v];
    return 1;
}

// ---------------- LSQ (lsq/factor_solve.ipp) ----------------

// lsq::factor for nDim = 2 (factor_solve.ipp:13-22): inverse of {x00, x01, x11}
template <class R> HC_HD void lsq_factor(const R* xm, R* xf)
{
    const R d = R(1.) / (xm[0] * xm[2] - xm[1] * xm[1]);
    xf[0] = d * xm[2];
    xf[1] = -d * xm[1];
    xf[2] = d * xm[0];
}

// lsq::solve for nDim = 2 (factor_solve.ipp:47-69): gradient along dir
template <class R> HC_HD R lsq_solve1(const R* xf, R q0, R q1, R d0, R d1)
{
    return (xf[0] * q0 + xf[1] * q1) * d0 + (xf[1] * q0 + xf[2] * q1) * d1;
}

}   // namespace hc
