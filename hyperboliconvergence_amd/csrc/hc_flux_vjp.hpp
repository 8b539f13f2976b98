// hc_flux_vjp.hpp -- hand-written reverse mode (vector-Jacobian products) of face fluxes.
//
// For a face flux F(q_L, q_R; minf) and a weight vector w (the difference of the residual
// adjoints of the face's two cells), the adjoint kernels need
//     q_L_bar = (dF/dq_L)^T w,   q_R_bar = (dF/dq_R)^T w,   minf_bar = (dF/dminf)^T w.
// Forward mode gets them from nine directional derivatives (Dual<9>); the reverse sweep below
// gets all nine from one forward evaluation and one backward pass over its intermediates.
// The forward part repeats the operations of hc_point.hpp (so the branches taken are the ones
// the flux takes), and every kink takes the branch the value takes, with the tie rules of
// hc_dual.hpp (fmax / fmin keep their first operand on ties, |x| has slope +1 at 0), so the
// products equal the forward-mode ones up to rounding.
//
// AUSM+up: euler/fluxes/ausmPlusUP.ipp:6-171 (with the Mach splittings of
// euler/fluxes/machSplittings.ipp:8-43 and the state transforms of
// euler/transforms/{conserved,primitive}.ipp).  The spectral radius (lambda) is not
// differentiated: the adjoint trajectories run with a fixed time step.
#pragma once

#include "hc_dual.hpp"

namespace hc {

// reverse of euler_from_conserved / euler_from_primitive: adds to qb the adjoint of the
// solution set given the adjoints of the state's u, v, r, p, h and a2 (inv_Rg = 1 / Rg; the
// reverse sweeps multiply by reciprocals instead of dividing, since they need not reproduce any
// forward rounding)
template <int BASIS>
HC_HD void euler_state_vjp(const Species<double>& sp, double inv_Rg, const double* q, const EState<double>& s,
                           double ub, double vb, double rb, double pb, double hb, double a2b, double* qb)
{
    const double gam = sp.gamma, Rg = sp.Rg, gam1 = sp.gamma1;
    // h = a2 * gam1 + k, a2 = gam * Rg * t, k = 0.5 * v2
    a2b = a2b + hb * gam1;
    const double kb = hb;
    const double tb = a2b * (gam * Rg);
    const double v2b = 0.5 * kb;
    const double inv_r = 1.0 / s.r;
    // t = p / (r Rg): dt = dp / (r Rg) - t dr / r
    pb = pb + tb * inv_r * inv_Rg;
    rb = rb - tb * s.t * inv_r;
    if constexpr (BASIS == BASIS_PRIMITIVE) {
        // v2 = u^2 + v^2
        qb[0] += ub + 2.0 * s.u * v2b;
        qb[1] += vb + 2.0 * s.v * v2b;
        qb[2] += rb;
        qb[3] += pb;
    } else {
        // p = (gam - 1)(re - r k); u = ru / r, v = rv / r
        const double k = 0.5 * s.v2;
        const double reb = pb * (gam - 1.0);
        rb = rb - pb * (gam - 1.0) * k;
        const double kb2 = -pb * (gam - 1.0) * s.r;
        const double v2b2 = v2b + 0.5 * kb2;
        ub = ub + 2.0 * s.u * v2b2;
        vb = vb + 2.0 * s.v * v2b2;
        qb[0] += ub * inv_r;
        qb[1] += vb * inv_r;
        const double r1b = ub * q[0] + vb * q[1];
        qb[2] += rb - r1b * inv_r * inv_r;
        qb[3] += reb;
    }
}

// Mach splittings and their slopes (machSplittings.ipp:8-43)
HC_HD double ms_m1_d(int sign, double M) { return 0.5 * (1.0 + double(sign) * (M < 0.0 ? -1.0 : 1.0)); }
HC_HD double ms_m2_d(int sign, double M) { return double(sign) * 0.5 * (M + double(sign)); }
HC_HD double ms_m4_d(int sign, double M, double beta)
{
    const double s16b = double(sign) * 16.0 * beta;
    return ms_m2_d(sign, M) * (1.0 - s16b * ms_m2(-sign, M)) - ms_m2(sign, M) * s16b * ms_m2_d(-sign, M);
}
HC_HD double ms_p1_d(int sign, double M)
{
    const double eps = NumLimits<double>::max();
    const double inv_b = 1.0 / (hc_fabs(M) + eps);
    const double r = (ms_m1(sign, M) + eps) * inv_b;
    return (ms_m1_d(sign, M) - r * (M < 0.0 ? -1.0 : 1.0)) * inv_b;
}
HC_HD double ms_p3_d(int sign, double M) { return ms_m2_d(sign, M) * (double(sign) * 2.0 - M) - ms_m2(sign, M); }
// d P5 / d M and d P5 / d alpha
HC_HD void ms_p5_d(int sign, double M, double alpha, double& dM, double& dalpha)
{
    const double s16 = double(sign) * 16.0;
    const double a = ms_m2(sign, M), b = ms_m2(-sign, M);
    const double da = ms_m2_d(sign, M), db = ms_m2_d(-sign, M);
    dM = ms_p3_d(sign, M) - s16 * alpha * (a * b + M * da * b + M * a * db);
    dalpha = -s16 * M * a * b;
}

// Limiter value and its partial derivatives in both arguments: the smooth van Albada limiter
// (limiters/limiter.ipp:27-38 without the sign switch) by hand with one reciprocal, the
// others from Dual<2> numbers.  The value has the forward path's bits.
HC_HD double limit_d(int lim, double a, double b, double& da, double& db)
{
    if (lim == LIM_SMOOTHVANALBADA) {
        const double eps = NumLimits<double>::min();
        const double num = b * (a * a) + a * (b * b);
        const double den = a * a + b * b + eps;
        const double f = num / den;
        const double inv = 1.0 / den;
        da = (2.0 * a * b + b * b - f * (2.0 * a)) * inv;
        db = (a * a + 2.0 * a * b - f * (2.0 * b)) * inv;
        return f;
    }
    Dual<2> x(a), y(b);
    x.d[0] = 1.0;
    y.d[1] = 1.0;
    const Dual<2> r = limit(lim, x, y);
    da = r.d[0];
    db = r.d[1];
    return r.v;
}

// AUSM+up face flux (euler_ausmpup on the states of ql, qr): adds (dF/dql)^T w to qlb,
// (dF/dqr)^T w to qrb and returns (dF/dminf)^T w
template <int BASIS>
HC_HD double ausmpup_vjp(const Species<double>& sp, int vscale, int pscale, const FaceG<double>& f, const double* ql,
                         const double* qr, const double* w, double* qlb, double* qrb)
{
    const EState<double> sl = euler_state<double, BASIS>(sp, ql);
    const EState<double> sr = euler_state<double, BASIS>(sp, qr);
    // ---- forward: the intermediates of euler_ausmpup ----
    const double alpha0 = 0.1875, beta = 0.125, Ku = 0.75 * 2.0, Kp = 0.25, sigma = 1.0;
    const double ascoeff = 2. * (sp.gamma - 1.) / (sp.gamma + 1.);
    const double unl = proj_vel(f.n0, f.n1, sl.u, sl.v);
    const double unr = proj_vel(f.n0, f.n1, sr.u, sr.v);
    const double als2 = ascoeff * sl.h;
    const double ars2 = ascoeff * sr.h;
    const double sqal = hc_sqrt(als2), sqar = hc_sqrt(ars2);
    const double dal = hc_fmax(unl, sqal), dar = hc_fmax(-unr, sqar);
    const bool dal_u = dal == unl, dar_u = dar == -unr;
    const double als = als2 / dal;
    const double ars = ars2 / dar;
    const double ra = 0.5 * (sl.r + sr.r);
    const double as = hc_fmin(ars, als);
    const bool as_r = as == ars;
    const double ml = unl / as;
    const double mr = unr / as;
    const double minf = sp.minf;
    const double m0 = hc_sqrt(0.5 * (ml * ml + mr * mr));
    const double mu = sp.lref / (as * sp.dt);
    // clamp(x, minf, 1): 0 = x, 1 = minf, 2 = 1
    auto clamp_branch = [&](double x) { return x < minf ? 1 : (1.0 < x ? 2 : 0); };
    const int bc = clamp_branch(m0);
    const double m0_c = std_clamp(m0, minf, 1.0);
    const double m0mu = hc_fmax(m0, mu);
    const bool mx_m0 = m0mu == m0;
    const int ba = clamp_branch(m0mu);
    const double m0_a = std_clamp(m0mu, minf, 1.0);
    const double m0_p = pscale == SCALE_CONVECTIVE ? m0_c : pscale == SCALE_ACOUSTIC ? 1.0 : m0_a;
    const double m0_u = vscale == SCALE_CONVECTIVE ? m0_c : vscale == SCALE_ACOUSTIC ? 1.0 : m0_a;
    const double fa_p = m0_p * (2. - m0_p);
    const double fa_u = m0_u * (2. - m0_u);
    const double mlsup = hc_fabs(ml) > 1.0 ? 1.0 : 0.0;
    const double mlsub = 1.0 - mlsup;
    const double mrsup = hc_fabs(mr) > 1.0 ? 1.0 : 0.0;
    const double mrsub = 1.0 - mrsup;
    const double mlp = mlsub * ms_m4(1, ml, beta) + mlsup * ms_m1(1, ml);
    const double mrm = mrsub * ms_m4(-1, mr, beta) + mrsup * ms_m1(-1, mr);
    const double delp = sr.p - sl.p;
    const double mpc0 = 1. - sigma * m0 * m0;
    const double mpc = hc_fmax(mpc0, 0.);
    const bool mpc_x = mpc == mpc0;
    const double D = fa_p * ra * as * as;
    const double mp = Kp * mpc * delp / D;
    const double S = mlp + mrm - mp;
    const double ua = as * S * f.area;
    const double alpha = alpha0 * (5. * fa_u * fa_u - 4.);
    const double plp = mlsub * ms_p5(1, ml, alpha) + mlsup * ms_p1(1, ml);
    const double prm = mrsub * ms_p5(-1, mr, alpha) + mrsup * ms_p1(-1, mr);
    const double delu = unr - unl;
    const double G = fa_u * as;
    const double m1l = (ua > 0.0) ? 1.0 : 0.0;
    const double m1r = (ua > 0.0) ? 0.0 : 1.0;
    const double rup = m1l * sl.r + m1r * sr.r;
    const double mdot = ua * rup;
    // ---- reverse ----
    const double inv_as = 1.0 / as;
    // F0 = mdot u_up + pa n0, F1 = mdot v_up + pa n1, F2 = mdot, F3 = mdot h_up
    const double mdotb = w[0] * (m1l * sl.u + m1r * sr.u) + w[1] * (m1l * sl.v + m1r * sr.v) + w[2] +
                         w[3] * (m1l * sl.h + m1r * sr.h);
    const double pab = w[0] * f.n0 + w[1] * f.n1;
    double ulb = (w[0] * mdot) * m1l, urb = (w[0] * mdot) * m1r;
    double vlb = (w[1] * mdot) * m1l, vrb = (w[1] * mdot) * m1r;
    double hlb = (w[3] * mdot) * m1l, hrb = (w[3] * mdot) * m1r;
    // mdot = ua * r_up
    const double uab = mdotb * rup;
    double rlb = mdotb * ua * m1l, rrb = mdotb * ua * m1r;
    // pa = (plp pl + prm pr - pu) A
    const double X = pab * f.area;
    double plpb = X * sl.p, prmb = X * sr.p;
    double plb = X * plp, prb = X * prm;
    const double pub = -X;
    // pu = Ku plp prm ra G delu, G = fa_u as
    const double kpp = Ku * plp * prm;
    plpb += pub * Ku * prm * ra * G * delu;
    prmb += pub * Ku * plp * ra * G * delu;
    double rab = pub * kpp * G * delu;
    const double Gb = pub * kpp * ra * delu;
    double fa_ub = Gb * as;
    double asb = Gb * fa_u;
    const double delub = pub * kpp * ra * G;
    double unrb = delub, unlb = -delub;
    // plp, prm: P5 below |M| = 1, P1 above
    double mlb = 0.0, mrb = 0.0, alphab = 0.0;
    {
        double dM, dA;
        ms_p5_d(1, ml, alpha, dM, dA);
        mlb += plpb * (mlsub * dM + mlsup * ms_p1_d(1, ml));
        alphab += plpb * mlsub * dA;
        ms_p5_d(-1, mr, alpha, dM, dA);
        mrb += prmb * (mrsub * dM + mrsup * ms_p1_d(-1, mr));
        alphab += prmb * mrsub * dA;
    }
    // alpha = alpha0 (5 fa_u^2 - 4)
    fa_ub += alphab * alpha0 * 10.0 * fa_u;
    // ua = as S A, S = mlp + mrm - mp
    asb += uab * S * f.area;
    const double Sb = uab * as * f.area;
    mlb += Sb * (mlsub * ms_m4_d(1, ml, beta) + mlsup * ms_m1_d(1, ml));
    mrb += Sb * (mrsub * ms_m4_d(-1, mr, beta) + mrsup * ms_m1_d(-1, mr));
    const double mpb = -Sb;
    // mp = Kp mpc delp / D, D = fa_p ra as as
    const double inv_D = 1.0 / D;
    const double mpcb = mpb * Kp * delp * inv_D;
    const double delpb = mpb * Kp * mpc * inv_D;
    const double Db = -mpb * mp * inv_D;
    double fa_pb = Db * ra * as * as;
    rab += Db * fa_p * as * as;
    asb += Db * 2.0 * fa_p * ra * as;
    double m0b = mpc_x ? mpcb * (-2.0 * sigma * m0) : 0.0;
    prb += delpb;
    plb -= delpb;
    // fa = m0 (2 - m0)
    double m0_pb = fa_pb * (2.0 - 2.0 * m0_p);
    double m0_ub = fa_ub * (2.0 - 2.0 * m0_u);
    double m0_cb = 0.0, m0_ab = 0.0;
    if (pscale == SCALE_CONVECTIVE) m0_cb += m0_pb;
    else if (pscale != SCALE_ACOUSTIC) m0_ab += m0_pb;
    if (vscale == SCALE_CONVECTIVE) m0_cb += m0_ub;
    else if (vscale != SCALE_ACOUSTIC) m0_ab += m0_ub;
    double minfb = 0.0;
    if (bc == 1) minfb += m0_cb;
    else if (bc == 0) m0b += m0_cb;
    if (ba == 1) minfb += m0_ab;
    else if (ba == 0) {
        if (mx_m0) m0b += m0_ab;
        else asb += m0_ab * (-mu * inv_as);   // mu = lref / (as dt)
    }
    // m0 = sqrt(0.5 (ml^2 + mr^2))
    const double Y = m0b * (0.5 / m0);
    mlb += Y * ml;
    mrb += Y * mr;
    // ml = unl / as, mr = unr / as
    unlb += mlb * inv_as;
    unrb += mrb * inv_as;
    asb -= (mlb * ml + mrb * mr) * inv_as;
    // as = fmin(ars, als)
    const double arsb = as_r ? asb : 0.0, alsb = as_r ? 0.0 : asb;
    // ra = (rl + rr) / 2
    rlb += 0.5 * rab;
    rrb += 0.5 * rab;
    // als = als2 / fmax(unl, sqrt(als2)), ars = ars2 / fmax(-unr, sqrt(ars2))
    const double inv_dal = 1.0 / dal;
    double als2b = alsb * inv_dal;
    const double dalb = -alsb * als * inv_dal;
    if (dal_u) unlb += dalb;
    else als2b += dalb * (0.5 / sqal);
    const double inv_dar = 1.0 / dar;
    double ars2b = arsb * inv_dar;
    const double darb = -arsb * ars * inv_dar;
    if (dar_u) unrb -= darb;
    else ars2b += darb * (0.5 / sqar);
    hlb += als2b * ascoeff;
    hrb += ars2b * ascoeff;
    // un = n0 u + n1 v
    ulb += unlb * f.n0;
    vlb += unlb * f.n1;
    urb += unrb * f.n0;
    vrb += unrb * f.n1;
    const double inv_Rg = 1.0 / sp.Rg;
    euler_state_vjp<BASIS>(sp, inv_Rg, ql, sl, ulb, vlb, rlb, plb, hlb, 0.0, qlb);
    euler_state_vjp<BASIS>(sp, inv_Rg, qr, sr, urb, vrb, rrb, prb, hrb, 0.0, qrb);
    return minfb;
}

}   // namespace hc
