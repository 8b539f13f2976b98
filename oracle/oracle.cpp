/*
 * oracle.cpp -- CPU restatement of the reference's structured-mesh finite-volume
 * residual + explicit Runge-Kutta hot path (lacrymose/hyperboliConvergence).
 *
 * TEST INFRASTRUCTURE ONLY: this file is the parity checker for the HIP product
 * path and the timed CPU baseline.  It is never linked into, loaded by, or
 * called from the product.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference's cpp/inc/).  Arithmetic is written in the reference's
 * evaluation order (left-associative, no re-association); the parity build uses
 * -O2 -ffp-contract=off and no fast-math so that every product, sum, division
 * and square root is one correctly rounded IEEE operation, exactly as the
 * reference's arithmetic is when built without -ffast-math.
 *
 * Field-level semantics that the reference leaves broken at HEAD (lsq/lsq.h,
 * spatial/muscl.h, spatial/residualCalc.h, boundary loops do not compile) are
 * fixed as follows:
 *  - MUSCL face states: q_L + d_L*lim(g_c, g_L.n), q_R - d_R*lim(g_c, g_R.n)
 *    with g_c = (q_R-q_L)/(d_L+d_R) (spatial/muscl.h:35-58, face.normal read as
 *    face.metric[0], ql/qr read as q_l/q_r).
 *  - LSQ neighbours are the four face neighbours; across a periodic boundary
 *    the neighbour is the periodic image (centre translated by the period
 *    vector); mirror points (spatial/lsqMetrics.h:104-163) are added to the
 *    X metric only on non-periodic sides; the Q metric has no boundary term on
 *    non-periodic sides (spatial/lsqMetrics.h:181-213 accumulates interior
 *    edges only).  Per cell the order is: interior i-edges, interior j-edges,
 *    boundary i side, boundary j side.
 *  - Periodic wrap faces are full MUSCL faces on the i=0 / j=0 node faces
 *    (spatial/boundary/periodicLoop.h:129-176) with the left cell replaced by
 *    its periodic image.
 *  - Ghost-cell faces are first order F(q_ghost, q_int) / F(q_int, q_ghost)
 *    (spatial/boundary/ghostCellLoop.h:133-224).
 *  - Inviscid wall flux uses q_wall = q_interior (dq = 0 in
 *    conservationLaws/euler/boundary/inviscidWall.ipp:54-57).
 *  - Local time stepping: dq = R_tot * ((beta_s*cfl)/lambda0_i), lambda0 being
 *    the stage-0 spectral radius of the cell (spatial/eulerForwardUpdate.h:155-160
 *    called with cfl -> beta_s*cfl, and the lambda of the stage-0 residual).
 */
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int MAXV = 4;

/* State<Euler,2> layout, conservationLaws/euler/euler.h:62-83 */
enum { SP = 0, SR = 1, ST = 2, SH = 3, SV2 = 4, SA2 = 5, SU = 6, SVV = 7 };
/* State<ScalarAdvection,2> layout, conservationLaws/scalarAdvection/scalarAdvection.h:48-59 */
enum { SQ = 0, SQU = 1, SQV = 2 };

struct Face {
    double area;
    double c[2];
    double n[2];   // metric[0]
    double t[2];   // metric[1]
};

struct Cell {
    double vol;
    double c[2];
};

struct FluxRes {
    double f[MAXV];
    double lambda;
};

inline int nvar_of(int law) { return law == ORC_LAW_EULER ? 4 : 3; }

/* ---------------- geometry (geometry/operations.ipp, surface.ipp, volume.ipp) ---------------- */

// geometry/operations.ipp:6-13
inline double dot2(const double* l, const double* r)
{
    double d = 0;
    d += l[0] * r[0];
    d += l[1] * r[1];
    return d;
}

// geometry/operations.ipp:21-25
inline double length2d(const double* d) { return std::sqrt(dot2(d, d)); }

// geometry/surface.ipp:12-29
Face surface(const double* p0, const double* p1)
{
    const double tangent0[2] = {p1[0] - p0[0], p1[1] - p0[1]};
    const double tangent1[2] = {p0[0] - p1[0], p0[1] - p1[1]};
    const double l = length2d(tangent0);
    const double l1 = 1. / l;
    Face f;
    f.area = l;
    f.c[0] = p0[0] + tangent0[0] * 0.5;
    f.c[1] = p0[1] + tangent0[1] * 0.5;
    const double s1[2] = {tangent1[0] * l1, tangent1[1] * l1};
    f.n[0] = -s1[1];   // orthog, operations.ipp:43-47
    f.n[1] = s1[0];
    f.t[0] = tangent0[0] * l1;
    f.t[1] = tangent0[1] * l1;
    return f;
}

// geometry/volume.ipp:25-66 with fem/fem.h:38-77 evaluated at the element centre
Cell volume(const double* p0, const double* p1, const double* p2, const double* p3)
{
    // fem::n2({0,0}) = {0.25,0.25,0.25,0.25}; fem::dn2({0,0}) rows:
    const double u[4] = {0.5 * 0.5, 0.5 * 0.5, 0.5 * 0.5, 0.5 * 0.5};
    const double du[4][2] = {{-0.5 * 0.5, 0.5 * -0.5},
                             {0.5 * 0.5, 0.5 * -0.5},
                             {-0.5 * 0.5, 0.5 * 0.5},
                             {0.5 * 0.5, 0.5 * 0.5}};
    Cell c;
    for (int i = 0; i < 2; ++i) {
        double com = 0.;
        com += u[0] * p0[i];
        com += u[1] * p1[i];
        com += u[2] * p2[i];
        com += u[3] * p3[i];
        c.c[i] = com;
    }
    double tangent[2][2] = {{0., 0.}, {0., 0.}};
    for (int i = 0; i < 2; ++i) {
        for (int j = 0; j < 2; ++j) {
            tangent[i][j] += du[0][i] * p0[j];
            tangent[i][j] += du[1][i] * p1[j];
            tangent[i][j] += du[2][i] * p2[j];
            tangent[i][j] += du[3][i] * p3[j];
        }
    }
    // cross, operations.ipp:27-32
    c.vol = std::fabs(tangent[0][0] * tangent[1][1] - tangent[0][1] * tangent[1][0]);
    return c;
}

/* ---------------- transforms (euler/transforms, scalarAdvection/transforms) ---------------- */

struct Species {
    double gamma, minf, lref, nu, pr, dt, R, gamma1;
};

inline Species to_species(const orc_species& s)
{
    return Species{s.gamma, s.minf, s.lref, s.nu, s.pr, s.dt, s.R, s.gamma1};
}

// euler/transforms/conserved.ipp:31-76
void euler_cons2state(const Species& sp, const double* qc, double* st)
{
    const double gam = sp.gamma;
    const double R = sp.R;
    const double gam1 = sp.gamma1;
    const double r = qc[2];
    const double re = qc[3];
    const double r1 = 1. / r;
    double v2 = 0;
    st[SU] = qc[0] * r1;
    st[SVV] = qc[1] * r1;
    v2 += st[SU] * st[SU];
    v2 += st[SVV] * st[SVV];
    st[SV2] = v2;
    const double k = 0.5 * v2;
    const double p = (gam - 1.) * (re - r * k);
    const double t = p / (R * r);
    const double a2 = gam * R * t;
    const double h = a2 * gam1 + k;
    st[SP] = p;
    st[SR] = r;
    st[ST] = t;
    st[SH] = h;
    st[SA2] = a2;
}

// euler/transforms/primitive.ipp:24-60
void euler_prim2state(const Species& sp, const double* qp, double* st)
{
    const double gam = sp.gamma;
    const double R = sp.R;
    const double gam1 = sp.gamma1;
    double v2 = 0;
    st[SU] = qp[0];
    v2 += qp[0] * qp[0];
    st[SVV] = qp[1];
    v2 += qp[1] * qp[1];
    st[SV2] = v2;
    const double k = 0.5 * v2;
    const double r = qp[2];
    const double p = qp[3];
    const double t = p / (r * R);
    const double a2 = gam * R * t;
    const double h = a2 * gam1 + k;
    st[SH] = h;
    st[SP] = p;
    st[SR] = r;
    st[ST] = t;
    st[SA2] = a2;
}

// euler/transforms/conserved.ipp:2-29
void euler_state2cons(const double* st, double* qc)
{
    const double r = st[SR];
    const double p = st[SP];
    const double h = st[SH];
    const double re = r * h - p;
    qc[0] = r * st[SU];
    qc[1] = r * st[SVV];
    qc[2] = r;
    qc[3] = re;
}

// euler/transforms/primitive.ipp:5-21
void euler_state2prim(const double* st, double* qp)
{
    qp[0] = st[SU];
    qp[1] = st[SVV];
    qp[2] = st[SR];
    qp[3] = st[SP];
}

// scalarAdvection/transforms/conserved.ipp:19-34 (set2State), :2-17 (state2Set)
void scalar_cons2state(const double* qc, double* st)
{
    st[SQU] = qc[0];
    st[SQV] = qc[1];
    st[SQ] = qc[2];
}
void scalar_state2cons(const double* st, double* qc)
{
    qc[0] = st[SQU];
    qc[1] = st[SQV];
    qc[2] = st[SQ];
}

// base/base.h:375-387 set2State dispatch on the solution basis
void set2state(const orc_case& cs, const Species& sp, const double* q, double* st)
{
    if (cs.law == ORC_LAW_SCALAR) {
        scalar_cons2state(q, st);
    } else if (cs.basis == ORC_BASIS_PRIMITIVE) {
        euler_prim2state(sp, q, st);
    } else {
        euler_cons2state(sp, q, st);
    }
}

void state2set(const orc_case& cs, int basis, const double* st, double* q)
{
    if (cs.law == ORC_LAW_SCALAR) {
        scalar_state2cons(st, q);
    } else if (basis == ORC_BASIS_PRIMITIVE) {
        euler_state2prim(st, q);
    } else {
        euler_state2cons(st, q);
    }
}

// base/base.h:250-273 set2Set<Conserved>(solution) and back
void sol2cons(const orc_case& cs, const Species& sp, const double* q, double* qc)
{
    const int nv = nvar_of(cs.law);
    if (cs.law == ORC_LAW_SCALAR || cs.basis == ORC_BASIS_CONSERVED) {
        for (int v = 0; v < nv; ++v) qc[v] = q[v];
        return;
    }
    double st[8];
    euler_prim2state(sp, q, st);
    euler_state2cons(st, qc);
}

void cons2sol(const orc_case& cs, const Species& sp, const double* qc, double* q)
{
    const int nv = nvar_of(cs.law);
    if (cs.law == ORC_LAW_SCALAR || cs.basis == ORC_BASIS_CONSERVED) {
        for (int v = 0; v < nv; ++v) q[v] = qc[v];
        return;
    }
    double st[8];
    euler_cons2state(sp, qc, st);
    euler_state2prim(st, q);
}

// euler/transforms/primitive.ipp:65-107 conserved delta -> primitive delta
void d_cons2prim(const Species& sp, const double* st, const double* dqc, double* dqp)
{
    const double dr = dqc[2];
    const double dre = dqc[3];
    const double rho1 = 1. / st[SR];
    const double k = 0.5 * st[SV2];
    double c = k * dr + dre;
    for (int i = 0; i < 2; i++) {
        const double u = st[SU + i];
        const double dru = dqc[i];
        const double du = (dru - u * dr) * rho1;
        dqp[i] = du;
        c -= u * dru;
    }
    const double dp = c * (sp.gamma - 1.);
    dqp[2] = dr;
    dqp[3] = dp;
}

// euler/transforms/primitive.ipp:110-156 primitive delta -> conserved delta
void d_prim2cons(const Species& sp, const double* st, const double* dqp, double* dqc)
{
    const double dr = dqp[2];
    const double dp = dqp[3];
    const double r = st[SR];
    double dk = 0;
    for (int i = 0; i < 2; i++) {
        const double u = st[SU + i];
        const double du = dqp[i];
        const double dru = r * du + u * dr;
        dqc[i] = dru;
        dk += u * du;
    }
    const double k = 0.5 * st[SV2];
    const double dre = dp * sp.gamma1 + r * dk + k * dr;
    dqc[2] = dr;
    dqc[3] = dre;
}

// euler/transforms/characteristic.ipp:4-41 characteristic delta -> primitive delta
void d_char2prim(const Species& sp, const double* st, const double* dqw, double* dqp)
{
    const double r = st[SR];
    const double a = std::sqrt(st[SA2]);
    const double dp = 0.5 * r * a * (dqw[2] - dqw[1]);
    const double dr = dp / (a * a) - r * dqw[0] / sp.gamma;
    dqp[0] = 0.5 * (dqw[1] + dqw[2]);
    dqp[1] = dqw[3];
    dqp[2] = dr;
    dqp[3] = dp;
}

// euler/transforms/characteristic.ipp:43-79 primitive delta -> characteristic delta
void d_prim2char(const Species& sp, const double* st, const double* dqp, double* dqw)
{
    const double r1 = 1. / st[SR];
    const double a1 = 1. / std::sqrt(st[SA2]);
    const double du = dqp[0];
    const double dr = dqp[2];
    const double dp = dqp[3];
    dqw[0] = sp.gamma * (dp * a1 * a1 - dr) * r1;
    dqw[1] = du - dp * r1 * a1;
    dqw[2] = du + dp * r1 * a1;
    dqw[3] = dqp[1];
}

/* ---------------- rotations (base/base.h:171-232) ---------------- */

// base/base.h:171-195, first nDim entries are the (single) vector quantity
inline void rotate_to(const Face& f, const double* q0, double* q1, int nv)
{
    for (int v = 0; v < nv; ++v) q1[v] = q0[v];
    double a = 0;
    a += f.n[0] * q0[0];
    a += f.n[1] * q0[1];
    double b = 0;
    b += f.t[0] * q0[0];
    b += f.t[1] * q0[1];
    q1[0] = a;
    q1[1] = b;
}

// base/base.h:201-209 (rotation by the transposed metric)
inline void rotate_from(const Face& f, const double* q0, double* q1, int nv)
{
    for (int v = 0; v < nv; ++v) q1[v] = q0[v];
    double a = 0;
    a += f.n[0] * q0[0];
    a += f.t[0] * q0[1];
    double b = 0;
    b += f.n[1] * q0[0];
    b += f.t[1] * q0[1];
    q1[0] = a;
    q1[1] = b;
}

// base/base.h:148-158
inline double projected_velocity(const double* n, const double* st, int uoff)
{
    double un = 0;
    un += n[0] * st[uoff];
    un += n[1] * st[uoff + 1];
    return un;
}

/* ---------------- fluxes ---------------- */

// euler/fluxes/exactFlux.ipp:4-36 (Euler) and scalarAdvection/fluxes/exactFlux.ipp:8-21
FluxRes exact_flux(const orc_case& cs, const double* n, const double* st)
{
    FluxRes fr{};
    if (cs.law == ORC_LAW_SCALAR) {
        const double un = projected_velocity(n, st, SQU);
        const double q = st[SQ];
        fr.f[0] = 0.;
        fr.f[1] = 0.;
        fr.f[2] = un * q;
        fr.lambda = std::abs(un);
        return fr;
    }
    const double r = st[SR];
    const double p = st[SP];
    const double h = st[SH];
    const double a = std::sqrt(st[SA2]);
    const double un = projected_velocity(n, st, SU);
    const double mn = r * un;
    for (int i = 0; i < 2; i++) fr.f[i] = mn * st[SU + i] + p * n[i];
    fr.f[2] = mn;
    fr.f[3] = mn * h;
    fr.lambda = std::abs(un) + a;
    return fr;
}

// euler/euler.h:161-166, scalarAdvection/scalarAdvection.h:93-98
inline double spectral_radius(const orc_case& cs, const double* n, const double* st)
{
    if (cs.law == ORC_LAW_SCALAR) return std::abs(projected_velocity(n, st, SQU));
    return std::abs(projected_velocity(n, st, SU)) + std::sqrt(st[SA2]);
}

// std::max as used by the reference (returns the first argument unless it is less)
inline double stdmax(double a, double b) { return (a < b) ? b : a; }

// base/fluxes/centralFlux.ipp:2-14 (as CentralFlux<Law>::flux, roeFlux.ipp:11-12)
FluxRes central_flux(const orc_case& cs, const Face& f, const double* sl, const double* sr)
{
    const int nv = nvar_of(cs.law);
    const FluxRes fl = exact_flux(cs, f.n, sl);
    const FluxRes fr = exact_flux(cs, f.n, sr);
    FluxRes out{};
    for (int v = 0; v < nv; ++v) out.f[v] = ((fl.f[v] + fr.f[v]) * 0.5) * f.area;
    out.lambda = stdmax(fl.lambda, fr.lambda) * f.area;
    return out;
}

void state2cons_law(const orc_case& cs, const double* st, double* qc)
{
    if (cs.law == ORC_LAW_SCALAR) scalar_state2cons(st, qc);
    else euler_state2cons(st, qc);
}

// base/fluxes/rusanovFlux.ipp:23-50
FluxRes rusanov_dissipation(const orc_case& cs, const Face& f, const double* sl, const double* sr)
{
    const int nv = nvar_of(cs.law);
    const double lambda = stdmax(spectral_radius(cs, f.n, sl), spectral_radius(cs, f.n, sr));
    double cr[MAXV], cl[MAXV];
    state2cons_law(cs, sr, cr);
    state2cons_law(cs, sl, cl);
    const double la = lambda * f.area;
    FluxRes out{};
    for (int v = 0; v < nv; ++v) out.f[v] = (cr[v] - cl[v]) * (-0.5 * la);
    out.lambda = lambda;
    return out;
}

// base/fluxes/rusanovFlux.ipp:2-20
FluxRes rusanov_flux(const orc_case& cs, const Face& f, const double* sl, const double* sr)
{
    const int nv = nvar_of(cs.law);
    FluxRes result = central_flux(cs, f, sl, sr);
    const FluxRes d = rusanov_dissipation(cs, f, sl, sr);
    for (int v = 0; v < nv; ++v) result.f[v] += d.f[v];
    return result;
}

// euler/fluxes/roe.ipp:2-48
void roe_average(const Species& sp, const double* sl, const double* sr, double* ravg)
{
    const double cr = std::sqrt(sr[SR] / sl[SR]);
    const double wl = 1. / (1. + cr);
    const double wr = 1. - wl;
    const double ra = cr * sl[SR];
    double ka = 0;
    for (int i = 0; i < 2; i++) {
        const double ua = wl * sl[SU + i] + wr * sr[SU + i];
        ka += ua * ua;
        ravg[SU + i] = ua;
    }
    ka *= 0.5;
    const double ha = wl * sl[SH] + wr * sr[SH];
    const double a2a = (sp.gamma - 1.) * (ha - ka);
    const double ta = a2a / (sp.R * sp.gamma);
    const double pa = ra * sp.R * ta;
    ravg[SP] = pa;
    ravg[SR] = ra;
    ravg[ST] = ta;
    ravg[SH] = ha;
    ravg[SV2] = 2. * ka;
    ravg[SA2] = a2a;
}

// euler/euler.h:168-181
void wavespeeds(const double* n, const double* st, double* w)
{
    const double u = projected_velocity(n, st, SU);
    const double a = std::sqrt(st[SA2]);
    w[0] = u;
    w[1] = u - a;
    w[2] = u + a;
    w[3] = u;
}

// euler/fluxes/roe.ipp:50-78
void entropyfix(const double* la, const double* ll, const double* lr, double* s)
{
    const double ua = std::fabs(la[0]);
    s[0] = ua;
    s[3] = ua;
    s[1] = std::fabs(la[1]);
    s[2] = std::fabs(la[2]);
    constexpr double eps = 0.05;
    for (int i = 1; i < 3; i++) {
        const double le = eps * std::fmax(0., std::fmax(la[i] - ll[i], lr[i] - la[i]));
        const double laa = std::fabs(la[i]);
        s[i] = (laa > le) ? laa : 0.5 * (le + laa * laa / le);
    }
}

// base/fluxes/roeFlux.ipp:22-58.  face_frame = true is the added variant whose Roe-average
// velocity is rotated into the face frame like the jump (the reference keeps it in the global frame).
FluxRes roe_dissipation(const orc_case& cs, const Species& sp, const Face& f, const double* sl, const double* sr,
                        bool face_frame = false)
{
    double ravg[8];
    roe_average(sp, sl, sr, ravg);
    double rot[8];
    for (int k = 0; k < 8; ++k) rot[k] = ravg[k];
    if (face_frame) {
        const double uv[2] = {ravg[SU], ravg[SVV]};
        double a[2];
        rotate_to(f, uv, a, 2);
        rot[SU] = a[0];
        rot[SVV] = a[1];
    }
    const double* rv = rot;
    double wa[4], wl[4], wr[4], lambda[4];
    wavespeeds(f.n, ravg, wa);
    wavespeeds(f.n, sl, wl);
    wavespeeds(f.n, sr, wr);
    entropyfix(wa, wl, wr, lambda);
    double cr[4], cl[4], dq[4], dqc[4];
    euler_state2cons(sr, cr);
    euler_state2cons(sl, cl);
    for (int v = 0; v < 4; ++v) dq[v] = cr[v] - cl[v];
    rotate_to(f, dq, dqc, 4);
    double dqp[4], dqw[4];
    d_cons2prim(sp, rv, dqc, dqp);
    d_prim2char(sp, rv, dqp, dqw);
    for (int v = 0; v < 4; ++v) dqw[v] = lambda[v] * dqw[v];
    double dqp2[4], dqc2[4], fd[4];
    d_char2prim(sp, rv, dqw, dqp2);
    d_prim2cons(sp, rv, dqp2, dqc2);
    rotate_from(f, dqc2, fd, 4);
    FluxRes out{};
    for (int v = 0; v < 4; ++v) out.f[v] = fd[v] * (-0.5 * f.area);
    out.lambda = spectral_radius(cs, f.n, ravg) * f.area;
    return out;
}

// base/fluxes/roeFlux.ipp:39-56
FluxRes roe_flux(const orc_case& cs, const Species& sp, const Face& f, const double* sl, const double* sr,
                 bool face_frame = false)
{
    const FluxRes c = central_flux(cs, f, sl, sr);
    const FluxRes d = roe_dissipation(cs, sp, f, sl, sr, face_frame);
    FluxRes out{};
    for (int v = 0; v < 4; ++v) out.f[v] = c.f[v] + d.f[v];
    out.lambda = std::fmax(c.lambda, d.lambda);
    return out;
}

// euler/fluxes/roeUnprecWS.ipp:1-109 (RoeUnprecWS, "Weiss & Smith / Li & Gu form", euler.h:392-405).
// Roe state of the two face states; flux = central - 0.5 (|u_n| dq + dU ruvh + dP onu) A, where
// dq is the conserved jump, ruvh = rho (u, v, 1, H) and onu = (n, 0, u_n) of the Roe state,
// dU = nu_p dp + nu_u du and dP = mu_p dp + mu_u du with the jumps of pressure and normal velocity
// and the coefficients of roeUnprecWS.ipp:61-69 (exponents A11 = 1, A12 = 0, A21 = 1, A22 = 1 on
// the species' free-stream Mach number, uinf = minf a); lambda = max(central, Roe-state spectral
// radius * area * minf^(A11 - 1)).
FluxRes roe_unprec_ws_flux(const orc_case& cs, const Species& sp, const Face& f, const double* sl, const double* sr)
{
    double ravg[8];
    roe_average(sp, sl, sr, ravg);
    const FluxRes central = central_flux(cs, f, sl, sr);
    const double dp = sr[SP] - sl[SP];
    const double du = projected_velocity(f.n, sr, SU) - projected_velocity(f.n, sl, SU);
    const double una = projected_velocity(f.n, ravg, SU);
    const double ra = ravg[SR];
    const double aa = std::sqrt(ravg[SA2]);
    const double minf = sp.minf;
    const double uinf = minf * aa;
    double cr[4], cl[4];
    euler_state2cons(sr, cr);
    euler_state2cons(sl, cl);
    const int A11 = 1, A12 = 0, A21 = 1, A22 = 1;
    const double nu_p = 1.00 * (1. / (ra * uinf * uinf)) * std::pow(minf, A11);
    const double nu_u = 0.05 * std::pow(minf, A12);
    const double mu_p = 0.02 * std::pow(minf, A21);
    const double mu_u = 0.10 * (ra * uinf) * std::pow(minf, A22);
    const double deltaU = nu_p * dp + nu_u * du;
    const double deltaP = mu_p * dp + mu_u * du;
    const double ruvh[4] = {ra * ravg[SU], ra * ravg[SU + 1], ra, ra * ravg[SH]};
    const double onu[4] = {f.n[0], f.n[1], 0., una};
    FluxRes out{};
    for (int v = 0; v < 4; ++v) {
        const double diff = 0.5 * (std::fabs(una) * (cr[v] - cl[v]) + deltaU * ruvh[v] + deltaP * onu[v]) * f.area;
        out.f[v] = central.f[v] - diff;
    }
    out.lambda = std::fmax(central.lambda, spectral_radius(cs, f.n, ravg) * f.area * std::pow(minf, A11 - 1));
    return out;
}

// euler/fluxes/machSplittings.ipp:8-43
inline double msM1(int sign, double M) { return 0.5 * (M + sign * std::fabs(M)); }
inline double msM2(int sign, double M) { return sign * 0.25 * (M + sign) * (M + sign); }
inline double msM4(int sign, double M, double beta) { return msM2(sign, M) * (1. - sign * 16. * beta * msM2(-sign, M)); }
inline double msP1(int sign, double M)
{
    constexpr double eps = std::numeric_limits<double>::max();
    return (msM1(sign, M) + eps) / (std::fabs(M) + eps);
}
inline double msP3(int sign, double M) { return msM2(sign, M) * (sign * 2. - M); }
inline double msP5(int sign, double M, double alpha)
{
    return msP3(sign, M) - sign * 16. * alpha * M * msM2(sign, M) * msM2(-sign, M);
}

inline double stdclamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// euler/fluxes/slau.ipp:4-161
FluxRes slau_flux(const orc_case& cs, const Species& sp, const Face& f, const double* sl, const double* sr)
{
    const double ua = std::sqrt(0.5 * (sl[SV2] + sr[SV2]));
    const double aa = 0.5 * (std::sqrt(sl[SA2]) + std::sqrt(sr[SA2]));
    const double aa1 = 1.0 / aa;
    const double ma = ua * aa1;
    const double ma_hat = std::fmin(1.0, ma);
    const double unl = projected_velocity(f.n, sl, SU);
    const double unr = projected_velocity(f.n, sr, SU);
    const double ml = unl * aa1;
    const double mr = unr * aa1;
    const double una = (sl[SR] * std::fabs(unl) + sr[SR] * std::fabs(unr)) / (sl[SR] + sr[SR]);
    const double chi = (1. - ma_hat) * (1. - ma_hat);
    // m0_c, m0_a and the scaling choices do not enter the SLAU/SLAU2 terms used (slau.ipp:37-64)
    const double g = -std::fmax(std::fmin(ml, 0.), -1.) * std::fmin(std::fmax(mr, 0.), 1.);
    const double dr = sr[SR] - sl[SR];
    const double dp = sr[SP] - sl[SP];
    const double mdot_c = sl[SR] * unl + sr[SR] * unr;
    const double mdot_dr = una * dr;
    const double mdot_dp = chi * dp * aa1;
    const double mdot = 0.5 * ((mdot_c - mdot_dr) * (1. - g) - mdot_dp) * f.area;
    const unsigned int mlsup = std::fabs(ml) > 1 ? 1 : 0;
    const unsigned int mlsub = 1 - mlsup;
    const unsigned int mrsup = std::fabs(mr) > 1 ? 1 : 0;
    const unsigned int mrsub = 1 - mrsup;
    const double betal = mlsub * msP3(1, ml) + mlsup * msP1(1, ml);
    const double betar = mrsub * msP3(-1, mr) + mrsup * msP1(-1, mr);
    const double p_c = (sl[SP] + sr[SP]);
    const double p_dp = (betal - betar) * dp;
    const double p_du = -ma * (betal + betar - 1.) * (p_c);
    const double pa = 0.5 * (p_c - p_dp - p_du) * f.area;
    const int m1l = (ua > 0) ? 1 : 0;
    const int m1r = (ua > 0) ? 0 : 1;
    FluxRes out{};
    for (int i = 0; i < 2; i++) out.f[i] = mdot * (m1l * sl[SU + i] + m1r * sr[SU + i]) + pa * f.n[i];
    out.f[2] = mdot;
    out.f[3] = mdot * (m1l * sl[SH] + m1r * sr[SH]);
    const double lmax = std::sqrt(std::fmax(sl[SA2], sr[SA2])) + std::fmax(std::fabs(unl), std::fabs(unr));
    out.lambda = lmax * f.area;
    return out;
}

// euler/fluxes/ausmPlusUP.ipp:6-171
FluxRes ausm_flux(const orc_case& cs, const Species& sp, const Face& f, const double* sl, const double* sr)
{
    const double alpha0 = 0.1875;
    const double beta = 0.125;
    const double Ku = 0.75 * 2.;
    const double Kp = 0.25;
    const double sigma = 1.0;
    const double ascoeff = 2. * (sp.gamma - 1.) / (sp.gamma + 1.);
    const double unl = projected_velocity(f.n, sl, SU);
    const double unr = projected_velocity(f.n, sr, SU);
    const double als2 = ascoeff * sl[SH];
    const double ars2 = ascoeff * sr[SH];
    const double als = als2 / std::fmax(unl, std::sqrt(als2));
    const double ars = ars2 / std::fmax(-unr, std::sqrt(ars2));
    const double ra = 0.5 * (sl[SR] + sr[SR]);
    const double as = std::fmin(ars, als);
    const double ml = unl / as;
    const double mr = unr / as;
    const double minf = sp.minf;
    const double m0 = std::sqrt(0.5 * (ml * ml + mr * mr));
    const double mu = sp.lref / (as * sp.dt);
    const double m0_c = stdclamp(m0, minf, 1.);
    const double m0_a = stdclamp(std::fmax(m0, mu), minf, 1.);
    const double m0_p = cs.pscale == ORC_SCALE_CONVECTIVE ? m0_c : cs.pscale == ORC_SCALE_ACOUSTIC ? 1.0 : m0_a;
    const double m0_u = cs.vscale == ORC_SCALE_CONVECTIVE ? m0_c : cs.vscale == ORC_SCALE_ACOUSTIC ? 1.0 : m0_a;
    const double fa_p = m0_p * (2. - m0_p);
    const double fa_u = m0_u * (2. - m0_u);
    const unsigned int mlsup = std::fabs(ml) > 1 ? 1 : 0;
    const unsigned int mlsub = 1 - mlsup;
    const unsigned int mrsup = std::fabs(mr) > 1 ? 1 : 0;
    const unsigned int mrsub = 1 - mrsup;
    const double mlp = mlsub * msM4(1, ml, beta) + mlsup * msM1(1, ml);
    const double mrm = mrsub * msM4(-1, mr, beta) + mrsup * msM1(-1, mr);
    const double delp = sr[SP] - sl[SP];
    const double mp = Kp * std::fmax(1. - sigma * m0 * m0, 0.) * delp / (fa_p * ra * as * as);
    const double ua = as * (mlp + mrm - mp) * f.area;
    const double alpha = alpha0 * (5. * fa_u * fa_u - 4.);
    const double plp = mlsub * msP5(1, ml, alpha) + mlsup * msP1(1, ml);
    const double prm = mrsub * msP5(-1, mr, alpha) + mrsup * msP1(-1, mr);
    const double delu = unr - unl;
    const double pu = Ku * plp * prm * ra * (fa_u * as) * delu;
    const double pa = (plp * sl[SP] + prm * sr[SP] - pu) * f.area;
    const int m1l = (ua > 0) ? 1 : 0;
    const int m1r = (ua > 0) ? 0 : 1;
    const double mdot = ua * (m1l * sl[SR] + m1r * sr[SR]);
    FluxRes out{};
    for (int i = 0; i < 2; i++) out.f[i] = mdot * (m1l * sl[SU + i] + m1r * sr[SU + i]) + pa * f.n[i];
    out.f[2] = mdot;
    out.f[3] = mdot * (m1l * sl[SH] + m1r * sr[SH]);
    const double lmd = 0.5 * (m0_p + 1) / fa_p;
    const double amax = std::sqrt(std::fmax(as * as, std::fmax(sl[SA2], sr[SA2])));
    const double lmax = lmd * amax + std::fmax(std::fabs(ua), std::fmax(std::fabs(unl), std::fabs(unr)));
    out.lambda = lmax * f.area;
    return out;
}

// base/base.h:375-387: FluxInterface converts both solution sets to States, then Flux::flux
FluxRes flux_states(const orc_case& cs, const Species& sp, const Face& f, const double* sl, const double* sr)
{
    switch (cs.flux) {
    case ORC_FLUX_CENTRAL: return central_flux(cs, f, sl, sr);
    case ORC_FLUX_RUSANOV: return rusanov_flux(cs, f, sl, sr);
    case ORC_FLUX_ROE: return roe_flux(cs, sp, f, sl, sr);
    case ORC_FLUX_SLAU: return slau_flux(cs, sp, f, sl, sr);
    case ORC_FLUX_ROE_FACEFRAME: return roe_flux(cs, sp, f, sl, sr, true);
    case ORC_FLUX_ROE_UNPREC_WS: return roe_unprec_ws_flux(cs, sp, f, sl, sr);
    default: return ausm_flux(cs, sp, f, sl, sr);
    }
}

FluxRes flux_sets(const orc_case& cs, const Species& sp, const Face& f, const double* ql, const double* qr)
{
    double sl[8], sr[8];
    set2state(cs, sp, ql, sl);
    set2state(cs, sp, qr, sr);
    return flux_states(cs, sp, f, sl, sr);
}

/* ---------------- limiters (limiters/limiter.ipp) ---------------- */

double limit(int lim, double a, double b)
{
    switch (lim) {
    case ORC_LIM_NOLIMIT1:   // limiter.ipp:143-147
        return 0;
    case ORC_LIM_NOLIMIT2:   // limiter.ipp:149-153
        return 0.5 * (a + b);
    case ORC_LIM_NOLIMIT3:   // limiter.ipp:155-159
        return 0.33333333333333333 * (2. * a + b);
    case ORC_LIM_VANALBADA2: {   // limiter.ipp:26-37
        const double eps = std::numeric_limits<double>::min();
        return a * b < 0 ? 0 : (b * (a * a) + a * (b * b)) / (a * a + b * b + eps);
    }
    case ORC_LIM_MINMOD2:   // limiter.ipp:39-47
        return a > 0 ? std::fmin(a, std::fmax(b, 0.)) : std::fmax(a, std::fmin(b, 0.));
    case ORC_LIM_SUPERBEE2:   // limiter.ipp:49-62
        return a > 0 ? std::fmax(0., std::fmax(std::fmin(2 * a, b), std::fmin(2 * b, a)))
                     : std::fmin(0., std::fmin(std::fmax(2 * a, b), std::fmax(2 * b, a)));
    case ORC_LIM_MC2:   // limiter.ipp:64-79
        return a > 0 ? std::fmax(0., std::fmin(0.5 * (a + b), std::fmin(2 * b, 2 * a)))
                     : std::fmin(0., std::fmax(0.5 * (a + b), std::fmax(2 * b, 2 * a)));
    case ORC_LIM_VANLEER2: {   // limiter.ipp:81-89
        const double eps = std::numeric_limits<double>::min();
        return (b + std::fabs(b)) / (a + std::fabs(b) + eps);
    }
    case ORC_LIM_CADA3: {   // limiter.ipp:91-161
        double t = 0.0;
        double gam = 1.5;
        double c1 = -2. / 7.;
        double c2 = 0.4;
        double c3 = 3 * gam - 2;
        double u, v, w, p, q;
        if (a > 0) { u = a; v = b; }
        else { u = -a; v = -b; }
        p = (v + 2. * u) / 3.;
        q = std::fabs(a) + std::fabs(b);
        if (q > t) {
            if (v < -2 * u) w = 0;
            else if (v < c1 * u) w = (v + 2. * u) / 3.;
            else if (v < 0) w = -2. * v;
            else if (v < c2 * u) w = 2. * v;
            else if (v < c3 * u) w = (v + 2. * u) / 3.;
            else w = gam * u;
        } else {
            w = p;
        }
        if (a <= 0) w = -w;
        return w;
    }
    default: {   // smooth van Albada: limiter.ipp:26-37 without the a*b<0 switch
        const double eps = std::numeric_limits<double>::min();
        return (b * (a * a) + a * (b * b)) / (a * a + b * b + eps);
    }
    }
}

/* ---------------- boundary point functions ---------------- */

// euler/fluxes/roe.ipp:2-48 applied to VariableSets (base/base.h:346-360)
// euler/boundary/riemann.ipp:18-113
int euler_riemann_update(const orc_case& cs, const Species& sp, const Face& f, const Cell& c,
                         const double* qint, const double* qfar, double* out)
{
    const double gam = sp.gamma;
    const double gam1 = sp.gamma1;
    const double into[2] = {c.c[0] - f.c[0], c.c[1] - f.c[1]};
    const int ok = dot2(into, f.n) > 0;
    double ri[4], rf[4], si[8], sf[8], sa[8];
    rotate_to(f, qint, ri, 4);
    rotate_to(f, qfar, rf, 4);
    set2state(cs, sp, ri, si);
    set2state(cs, sp, rf, sf);
    roe_average(sp, si, sf, sa);
    if ((sa[SU] / std::sqrt(sa[SV2])) < 0.001) {
        for (int v = 0; v < 4; ++v) out[v] = qint[v];
        return ok;
    }
    const double ma = sa[SU] / std::sqrt(sa[SA2]);
    const bool inflow = ma > 0;
    const double* quw = inflow ? qfar : qint;
    if (std::abs(ma) > 1.) {
        for (int v = 0; v < 4; ++v) out[v] = quw[v];
        return ok;
    }
    const double* stI = si;   // set2State(rotateToFace(face, qinterior))
    const double* stF = sf;
    const double ai = std::sqrt(stI[SA2]);
    const double af = std::sqrt(stF[SA2]);
    const double uni = stI[SU];
    const double unf = stF[SU];
    const double rim = uni - 2. * ai * gam1;
    const double rip = unf + 2. * af * gam1;
    const double ub = 0.5 * (rim + rip);
    const double ab = 0.25 * (gam - 1.) * (rip - rim);
    double ru[4], su[8];
    rotate_to(f, quw, ru, 4);
    set2state(cs, sp, ru, su);
    const double sb = su[SA2] / (gam * std::pow(su[SR], gam - 1.));
    const double vb = su[SVV];
    const double rb = std::pow(ab * ab / (gam * sb), gam1);
    const double pb = rb * ab * ab / gam;
    const double qb[4] = {ub, vb, rb, pb};
    double qs[4];
    if (cs.basis == ORC_BASIS_PRIMITIVE) {
        for (int v = 0; v < 4; ++v) qs[v] = qb[v];
    } else {
        double st[8];
        euler_prim2state(sp, qb, st);
        euler_state2cons(st, qs);
    }
    rotate_from(f, qs, out, 4);
    return ok;
}

// euler/boundary/entropy.ipp:18-88
int euler_entropy_update(const orc_case& cs, const Species& sp, const Face& f, const Cell& c,
                         const double* qint, const double* qfar, double* out)
{
    const double gam = sp.gamma;
    const double into[2] = {c.c[0] - f.c[0], c.c[1] - f.c[1]};
    const int ok = dot2(into, f.n) > 0;
    double ri[4], rf[4], si[8], sf[8], sa[8];
    rotate_to(f, qint, ri, 4);
    rotate_to(f, qfar, rf, 4);
    set2state(cs, sp, ri, si);
    set2state(cs, sp, rf, sf);
    roe_average(sp, si, sf, sa);
    if ((sa[SU] / std::sqrt(sa[SV2])) < 0.0001) {
        for (int v = 0; v < 4; ++v) out[v] = qint[v];
        return ok;
    }
    const double ma = sa[SU] / std::sqrt(sa[SA2]);
    const bool inflow = ma > 0;
    const double* quw = inflow ? qfar : qint;
    const double* qdw = inflow ? qint : qfar;
    if (std::abs(ma) > 1.) {
        for (int v = 0; v < 4; ++v) out[v] = quw[v];
        return ok;
    }
    double ru[4], rd[4], su[8], sd[8];
    rotate_to(f, quw, ru, 4);
    rotate_to(f, qdw, rd, 4);
    set2state(cs, sp, ru, su);
    set2state(cs, sp, rd, sd);
    const double pb = sd[SP];
    const double sb = su[SA2] / (gam * std::pow(su[SR], gam - 1.));
    const double rb = std::pow(pb / sb, 1. / gam);
    const double qb[4] = {su[SU], su[SVV], rb, pb};
    double qs[4];
    if (cs.basis == ORC_BASIS_PRIMITIVE) {
        for (int v = 0; v < 4; ++v) qs[v] = qb[v];
    } else {
        double st[8];
        euler_prim2state(sp, qb, st);
        euler_state2cons(st, qs);
    }
    rotate_from(f, qs, out, 4);
    return ok;
}

// scalarAdvection/boundary/riemann.ipp:15-40
int scalar_riemann_update(const orc_case& cs, const Face& f, const double* qint, const double* qref, double* out)
{
    double avg[3], rot[3], st[8];
    for (int v = 0; v < 3; ++v) avg[v] = qint[v] + (qref[v] - qint[v]) * 0.5;
    rotate_to(f, avg, rot, 3);
    scalar_cons2state(rot, st);
    const bool inflow = projected_velocity(f.n, st, SQU) > 0;
    for (int v = 0; v < 3; ++v) out[v] = inflow ? qref[v] : qint[v];
    return 1;
}

// spatial/boundary/boundaryCondition.h:63-80 (Fixed), euler/boundary/inviscidWall.ipp:15-28 (wall update)
int bc_update(const orc_case& cs, const Species& sp, int type, const Face& f, const Cell& c,
              const double* qint, const double* qref, double* out)
{
    const int nv = nvar_of(cs.law);
    if (type == ORC_BC_RIEMANN) {
        if (cs.law == ORC_LAW_SCALAR) return scalar_riemann_update(cs, f, qint, qref, out);
        return euler_riemann_update(cs, sp, f, c, qint, qref, out);
    }
    if (type == ORC_BC_ENTROPY) return euler_entropy_update(cs, sp, f, c, qint, qref, out);
    for (int v = 0; v < nv; ++v) out[v] = qref[v];
    return 1;
}

// euler/boundary/inviscidWall.ipp:33-72 with dq = 0
FluxRes wall_flux(const orc_case& cs, const Species& sp, const Face& f, const double* qint)
{
    double st[8];
    set2state(cs, sp, qint, st);
    const double pressureWall = st[SP];
    FluxRes fr{};
    for (int i = 0; i < 2; ++i) fr.f[i] = f.area * pressureWall * f.n[i];
    fr.f[2] = 0.;
    fr.f[3] = 0.;
    fr.lambda = f.area * spectral_radius(cs, f.n, st);
    return fr;
}

/* ---------------- least squares (lsq/xmetric.ipp, qmetric.ipp, factor_solve.ipp) ---------------- */

struct XM { double x[3]; };

// lsq/xmetric.ipp:53-69
inline XM xmetric(const double* dx)
{
    XM xm;
    xm.x[0] = dx[0] * dx[0];
    xm.x[1] = dx[0] * dx[1];
    xm.x[2] = dx[1] * dx[1];
    return xm;
}

// lsq/factor_solve.ipp:13-22 then :47-69
void lsq_solve(const double* xm, const double* qm, int nv, const double* dir, double* dq)
{
    const double d = 1. / (xm[0] * xm[2] - xm[1] * xm[1]);
    const double f0 = d * xm[2];
    const double f1 = -d * xm[1];
    const double f2 = d * xm[0];
    for (int v = 0; v < nv; ++v) {
        const double q0 = qm[2 * v + 0];
        const double q1 = qm[2 * v + 1];
        dq[v] = (f0 * q0 + f1 * q1) * dir[0] + (f1 * q0 + f2 * q1) * dir[1];
    }
}

/* ---------------- field level ---------------- */

struct Mesh {
    int ni, nj;
    const double* nodes;
    std::vector<Cell> cells;
    const double* node(int i, int j) const { return nodes + 2 * ((size_t)i * (nj + 1) + j); }
    const Cell& cell(int i, int j) const { return cells[(size_t)i * nj + j]; }
};

void build_cells(Mesh& m)
{
    m.cells.resize((size_t)m.ni * m.nj);
    // mesh/dual.ipp:42-83, corners o0{0,0} o1{1,0} o2{0,1} o3{1,1}
    for (int i = 0; i < m.ni; ++i)
        for (int j = 0; j < m.nj; ++j)
            m.cells[(size_t)i * m.nj + j] = volume(m.node(i, j), m.node(i + 1, j), m.node(i, j + 1), m.node(i + 1, j + 1));
}

inline Cell image_of(const Cell& c, const double* period)
{
    Cell im = c;
    im.c[0] = c.c[0] - period[0];
    im.c[1] = c.c[1] - period[1];
    return im;
}

// spatial/lsqMetrics.h:31-62, 104-163 and parallalg/neighbour_algorithm.h:253-317
void xmetrics(const orc_case& cs, const Mesh& m, std::vector<XM>& X)
{
    const int ni = m.ni, nj = m.nj;
    X.assign((size_t)ni * nj, XM{{0., 0., 0.}});
    auto acc = [&](int i, int j, const XM& a) {
        XM& x = X[(size_t)i * nj + j];
        for (int k = 0; k < 3; ++k) x.x[k] = x.x[k] + a.x[k];
    };
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni - 1; ++i) {
            const double dx[2] = {m.cell(i + 1, j).c[0] - m.cell(i, j).c[0], m.cell(i + 1, j).c[1] - m.cell(i, j).c[1]};
            const XM a = xmetric(dx);
            acc(i, j, a);
            acc(i + 1, j, a);
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj - 1; ++j) {
            const double dx[2] = {m.cell(i, j + 1).c[0] - m.cell(i, j).c[0], m.cell(i, j + 1).c[1] - m.cell(i, j).c[1]};
            const XM a = xmetric(dx);
            acc(i, j, a);
            acc(i, j + 1, a);
        }
    auto mirror = [&](const Cell& c0, const Cell& c1) {
        const double disp[2] = {c0.c[0] - c1.c[0], c0.c[1] - c1.c[1]};
        const double bp[2] = {c0.c[0] + disp[0], c0.c[1] + disp[1]};
        const double dx[2] = {bp[0] - c0.c[0], bp[1] - c0.c[1]};
        return xmetric(dx);
    };
    for (int j = 0; j < nj; ++j) {
        if (cs.bc[0] == ORC_BC_PERIODIC) {
            const Cell l = image_of(m.cell(ni - 1, j), cs.period_i);
            const Cell& r = m.cell(0, j);
            const double dx[2] = {r.c[0] - l.c[0], r.c[1] - l.c[1]};
            const XM a = xmetric(dx);
            acc(ni - 1, j, a);
            acc(0, j, a);
        } else {
            acc(0, j, mirror(m.cell(0, j), m.cell(1, j)));
        }
        if (cs.bc[1] != ORC_BC_PERIODIC) acc(ni - 1, j, mirror(m.cell(ni - 1, j), m.cell(ni - 2, j)));
    }
    for (int i = 0; i < ni; ++i) {
        if (cs.bc[2] == ORC_BC_PERIODIC) {
            const Cell l = image_of(m.cell(i, nj - 1), cs.period_j);
            const Cell& r = m.cell(i, 0);
            const double dx[2] = {r.c[0] - l.c[0], r.c[1] - l.c[1]};
            const XM a = xmetric(dx);
            acc(i, nj - 1, a);
            acc(i, 0, a);
        } else {
            acc(i, 0, mirror(m.cell(i, 0), m.cell(i, 1)));
        }
        if (cs.bc[3] != ORC_BC_PERIODIC) acc(i, nj - 1, mirror(m.cell(i, nj - 1), m.cell(i, nj - 2)));
    }
}

// spatial/lsqMetrics.h:181-213, lsq/qmetric.ipp:48-70; QMetric stored [nvar][2]
void qmetrics(const orc_case& cs, const Mesh& m, const double* q, std::vector<double>& Q)
{
    const int ni = m.ni, nj = m.nj, nv = nvar_of(cs.law);
    const int qs = 2 * nv;
    Q.assign((size_t)ni * nj * qs, 0.);
    auto edge = [&](int il, int jl, int ir, int jr, const double* cl, const double* cr) {
        const double dx[2] = {cr[0] - cl[0], cr[1] - cl[1]};
        const double* ql = q + ((size_t)il * nj + jl) * nv;
        const double* qr = q + ((size_t)ir * nj + jr) * nv;
        double* Ql = &Q[((size_t)il * nj + jl) * qs];
        double* Qr = &Q[((size_t)ir * nj + jr) * qs];
        for (int v = 0; v < nv; ++v) {
            const double dq = qr[v] - ql[v];
            for (int d = 0; d < 2; ++d) {
                const double a = dq * dx[d];
                Ql[2 * v + d] = Ql[2 * v + d] + a;
                Qr[2 * v + d] = Qr[2 * v + d] + a;
            }
        }
    };
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni - 1; ++i) edge(i, j, i + 1, j, m.cell(i, j).c, m.cell(i + 1, j).c);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj - 1; ++j) edge(i, j, i, j + 1, m.cell(i, j).c, m.cell(i, j + 1).c);
    if (cs.bc[0] == ORC_BC_PERIODIC)
        for (int j = 0; j < nj; ++j) {
            const Cell l = image_of(m.cell(ni - 1, j), cs.period_i);
            edge(ni - 1, j, 0, j, l.c, m.cell(0, j).c);
        }
    if (cs.bc[2] == ORC_BC_PERIODIC)
        for (int i = 0; i < ni; ++i) {
            const Cell l = image_of(m.cell(i, nj - 1), cs.period_j);
            edge(i, nj - 1, i, 0, l.c, m.cell(i, 0).c);
        }
}

// spatial/muscl.h:10-60 (general) and :62-87 (NoLimit1 overload)
FluxRes muscl_flux(const orc_case& cs, const Species& sp, const Face& f, const Cell& cl, const Cell& cr,
                   const double* ql, const double* qr, const double* xl, const double* xr,
                   const double* Ql, const double* Qr)
{
    const int nv = nvar_of(cs.law);
    if (cs.limiter == ORC_LIM_NOLIMIT1) return flux_sets(cs, sp, f, ql, qr);
    const double dx_c[2] = {cr.c[0] - cl.c[0], cr.c[1] - cl.c[1]};
    double dq_c[MAXV];
    for (int v = 0; v < nv; ++v) dq_c[v] = qr[v] - ql[v];
    const double tl[2] = {f.c[0] - cl.c[0], f.c[1] - cl.c[1]};
    const double tr[2] = {cr.c[0] - f.c[0], cr.c[1] - f.c[1]};
    const double dl = length2d(tl);
    const double dr = length2d(tr);
    const double s1 = 1. / (dl + dr);
    double dqdx_c[MAXV];
    for (int v = 0; v < nv; ++v) dqdx_c[v] = dq_c[v] * s1;
    // lsq::bias_solves, lsq/factor_solve.ipp:71-100
    const XM dxm_c = xmetric(dx_c);
    double xbl[3], xbr[3];
    for (int k = 0; k < 3; ++k) {
        xbl[k] = xl[k] - dxm_c.x[k];
        xbr[k] = xr[k] - dxm_c.x[k];
    }
    double qbl[2 * MAXV], qbr[2 * MAXV];
    for (int v = 0; v < nv; ++v)
        for (int d = 0; d < 2; ++d) {
            const double c = dq_c[v] * dx_c[d];
            qbl[2 * v + d] = Ql[2 * v + d] - c;
            qbr[2 * v + d] = Qr[2 * v + d] - c;
        }
    double gl[MAXV], gr[MAXV];
    lsq_solve(xbl, qbl, nv, f.n, gl);
    lsq_solve(xbr, qbr, nv, f.n, gr);
    double qlf[MAXV], qrf[MAXV];
    for (int v = 0; v < nv; ++v) {
        const double slope_l = limit(cs.limiter, dqdx_c[v], gl[v]);
        const double slope_r = limit(cs.limiter, dqdx_c[v], gr[v]);
        qlf[v] = ql[v] + slope_l * dl;
        qrf[v] = qr[v] - slope_r * dr;
    }
    return flux_sets(cs, sp, f, qlf, qrf);
}

inline void acc_plus(FluxRes& r, const FluxRes& f, int nv)   // base/fluxResult.ipp:8-14
{
    for (int v = 0; v < nv; ++v) r.f[v] += f.f[v];
    r.lambda += f.lambda;
}
inline void acc_minus(FluxRes& r, const FluxRes& f, int nv)   // base/fluxResult.ipp:16-22
{
    for (int v = 0; v < nv; ++v) r.f[v] -= f.f[v];
    r.lambda += f.lambda;
}

struct Field {
    const orc_case* cs;
    Species sp;
    Mesh mesh;
    int nv;
    std::vector<XM> X;
    std::vector<double> Q;
    std::vector<double> ghost[4];
    const double* bref;
    int bc_ok;
};

const double* bref_ptr(const Field& F, int b, int k)
{
    const int ni = F.mesh.ni, nj = F.mesh.nj;
    const size_t off = b == 0 ? 0 : b == 1 ? nj : b == 2 ? 2 * (size_t)nj : 2 * (size_t)nj + ni;
    return F.bref + (off + k) * F.nv;
}

// spatial/boundary/boundaryUpdate.h:137-266 (faces into the domain)
void boundary_update(Field& F, const double* q)
{
    const Mesh& m = F.mesh;
    const int ni = m.ni, nj = m.nj, nv = F.nv;
    const orc_case& cs = *F.cs;
    for (int b = 0; b < 4; ++b) {
        const int n_along = b < 2 ? nj : ni;
        F.ghost[b].assign((size_t)n_along * nv, 0.);
        if (cs.bc[b] == ORC_BC_PERIODIC) continue;
        for (int k = 0; k < n_along; ++k) {
            int i, j;
            Face f;
            if (b == 0) { i = 0; j = k; f = surface(m.node(i, j), m.node(i, j + 1)); }
            else if (b == 1) { i = ni - 1; j = k; f = surface(m.node(i + 1, j + 1), m.node(i + 1, j)); }
            else if (b == 2) { i = k; j = 0; f = surface(m.node(i + 1, j), m.node(i, j)); }
            else { i = k; j = nj - 1; f = surface(m.node(i, j + 1), m.node(i + 1, j + 1)); }
            const double* qi = q + ((size_t)i * nj + j) * nv;
            const int ok = bc_update(cs, F.sp, cs.bc[b], f, m.cell(i, j), qi, bref_ptr(F, b, k), &F.ghost[b][(size_t)k * nv]);
            if (!ok) F.bc_ok = 0;
        }
    }
}

// spatial/residualCalc.h:27-69, 146-256, 279-307
void residual_calc(Field& F, const double* q, std::vector<FluxRes>& res)
{
    const Mesh& m = F.mesh;
    const int ni = m.ni, nj = m.nj, nv = F.nv;
    const orc_case& cs = *F.cs;
    const Species& sp = F.sp;
    const size_t qs = 2 * nv;
    res.assign((size_t)ni * nj, FluxRes{});
    auto Rc = [&](int i, int j) -> FluxRes& { return res[(size_t)i * nj + j]; };
    auto qc = [&](int i, int j) { return q + ((size_t)i * nj + j) * nv; };
    auto Xc = [&](int i, int j) { return F.X[(size_t)i * nj + j].x; };
    auto Qc = [&](int i, int j) { return &F.Q[((size_t)i * nj + j) * qs]; };
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni - 1; ++i) {
            const Face f = surface(m.node(i + 1, j), m.node(i + 1, j + 1));
            const FluxRes fr = muscl_flux(cs, sp, f, m.cell(i, j), m.cell(i + 1, j), qc(i, j), qc(i + 1, j),
                                          Xc(i, j), Xc(i + 1, j), Qc(i, j), Qc(i + 1, j));
            acc_minus(Rc(i, j), fr, nv);
            acc_plus(Rc(i + 1, j), fr, nv);
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj - 1; ++j) {
            const Face f = surface(m.node(i + 1, j + 1), m.node(i, j + 1));
            const FluxRes fr = muscl_flux(cs, sp, f, m.cell(i, j), m.cell(i, j + 1), qc(i, j), qc(i, j + 1),
                                          Xc(i, j), Xc(i, j + 1), Qc(i, j), Qc(i, j + 1));
            acc_minus(Rc(i, j), fr, nv);
            acc_plus(Rc(i, j + 1), fr, nv);
        }
    // boundaries in id order (residualCalc.h:301-305)
    for (int b = 0; b < 4; ++b) {
        const int type = cs.bc[b];
        if (type == ORC_BC_PERIODIC) {
            if (b == 0) {   // periodicLoop.h:129-152
                for (int j = 0; j < nj; ++j) {
                    const Face f = surface(m.node(0, j), m.node(0, j + 1));
                    const Cell l = image_of(m.cell(ni - 1, j), cs.period_i);
                    const FluxRes fr = muscl_flux(cs, sp, f, l, m.cell(0, j), qc(ni - 1, j), qc(0, j),
                                                  Xc(ni - 1, j), Xc(0, j), Qc(ni - 1, j), Qc(0, j));
                    acc_minus(Rc(ni - 1, j), fr, nv);
                    acc_plus(Rc(0, j), fr, nv);
                }
            } else if (b == 2) {   // periodicLoop.h:153-176
                for (int i = 0; i < ni; ++i) {
                    const Face f = surface(m.node(i + 1, 0), m.node(i, 0));
                    const Cell l = image_of(m.cell(i, nj - 1), cs.period_j);
                    const FluxRes fr = muscl_flux(cs, sp, f, l, m.cell(i, 0), qc(i, nj - 1), qc(i, 0),
                                                  Xc(i, nj - 1), Xc(i, 0), Qc(i, nj - 1), Qc(i, 0));
                    acc_minus(Rc(i, nj - 1), fr, nv);
                    acc_plus(Rc(i, 0), fr, nv);
                }
            }
            continue;
        }
        const int n_along = b < 2 ? nj : ni;
        for (int k = 0; k < n_along; ++k) {
            const double* g = &F.ghost[b][(size_t)k * nv];
            if (type == ORC_BC_WALL) {   // boundaryFluxLoop.h:150-261
                int i, j;
                Face f;
                if (b == 0) { i = 0; j = k; f = surface(m.node(i, j), m.node(i, j + 1)); }
                else if (b == 1) { i = ni - 1; j = k; f = surface(m.node(i + 1, j + 1), m.node(i + 1, j)); }
                else if (b == 2) { i = k; j = 0; f = surface(m.node(i + 1, j), m.node(i, j)); }
                else { i = k; j = nj - 1; f = surface(m.node(i, j + 1), m.node(i + 1, j + 1)); }
                acc_plus(Rc(i, j), wall_flux(cs, sp, f, qc(i, j)), nv);
            } else {   // ghostCellLoop.h:133-224
                if (b == 0) {
                    const Face f = surface(m.node(0, k), m.node(0, k + 1));
                    acc_plus(Rc(0, k), flux_sets(cs, sp, f, g, qc(0, k)), nv);
                } else if (b == 1) {
                    const Face f = surface(m.node(ni, k), m.node(ni, k + 1));
                    acc_minus(Rc(ni - 1, k), flux_sets(cs, sp, f, qc(ni - 1, k), g), nv);
                } else if (b == 2) {
                    const Face f = surface(m.node(k + 1, 0), m.node(k, 0));
                    acc_plus(Rc(k, 0), flux_sets(cs, sp, f, g, qc(k, 0)), nv);
                } else {
                    const Face f = surface(m.node(k + 1, nj), m.node(k, nj));
                    acc_minus(Rc(k, nj - 1), flux_sets(cs, sp, f, qc(k, nj - 1), g), nv);
                }
            }
        }
    }
}

// spatial/spectralRadius.h:19-48 (transform_reduce max, init numeric_limits::min())
double spectral_radius_field(const Mesh& m, const std::vector<FluxRes>& res)
{
    double r = std::numeric_limits<double>::min();
    const size_t n = (size_t)m.ni * m.nj;
    for (size_t c = 0; c < n; ++c) r = stdmax(r, res[c].lambda / m.cells[c].vol);
    return r;
}

bool valid_case(const orc_case& cs, int ni, int nj)
{
    if (ni < 3 || nj < 3) return false;
    if ((cs.bc[0] == ORC_BC_PERIODIC) != (cs.bc[1] == ORC_BC_PERIODIC)) return false;
    if ((cs.bc[2] == ORC_BC_PERIODIC) != (cs.bc[3] == ORC_BC_PERIODIC)) return false;
    if (cs.law == ORC_LAW_SCALAR && cs.flux > ORC_FLUX_RUSANOV) return false;
    if (cs.flux < 0 || cs.flux > ORC_FLUX_ROE_UNPREC_WS) return false;
    if (cs.nstages < 1 || cs.nstages > 6) return false;
    if (cs.dt_mode < ORC_DT_GLOBAL || cs.dt_mode > ORC_DT_FIXED) return false;
    return true;
}

void setup_field(Field& F, const orc_case* cs, int ni, int nj, const double* nodes, const double* bref)
{
    F.cs = cs;
    F.sp = to_species(cs->sp);
    F.mesh.ni = ni;
    F.mesh.nj = nj;
    F.mesh.nodes = nodes;
    build_cells(F.mesh);
    F.nv = nvar_of(cs->law);
    F.bref = bref;
    F.bc_ok = 1;
    xmetrics(*cs, F.mesh, F.X);
}

}   // namespace

/* ================================ C interface ================================ */

extern "C" {

int orc_nvar(int law) { return nvar_of(law); }

void orc_surface(const double* p0, const double* p1, orc_face* out)
{
    const Face f = surface(p0, p1);
    *out = orc_face{f.area, f.c[0], f.c[1], f.n[0], f.n[1], f.t[0], f.t[1]};
}

void orc_volume(const double* p0, const double* p1, const double* p2, const double* p3, double* out3)
{
    const Cell c = volume(p0, p1, p2, p3);
    out3[0] = c.vol;
    out3[1] = c.c[0];
    out3[2] = c.c[1];
}

void orc_dual(int ni, int nj, const double* nodes, double* cells)
{
    Mesh m;
    m.ni = ni;
    m.nj = nj;
    m.nodes = nodes;
    build_cells(m);
    for (size_t c = 0; c < (size_t)ni * nj; ++c) {
        cells[3 * c] = m.cells[c].vol;
        cells[3 * c + 1] = m.cells[c].c[0];
        cells[3 * c + 2] = m.cells[c].c[1];
    }
}

// mesh/generate/twoD.h:15-40
void orc_mesh_linspace(int ni, int nj, double lox, double hix, double loy, double hiy, double* nodes)
{
    const double dx = (hix - lox) / (size_t)ni;
    const double dy = (hiy - loy) / (size_t)nj;
    for (size_t i = 0; i <= (size_t)ni; ++i)
        for (size_t j = 0; j <= (size_t)nj; ++j) {
            nodes[2 * (i * (nj + 1) + j)] = lox + i * dx;
            nodes[2 * (i * (nj + 1) + j) + 1] = loy + j * dy;
        }
}

// mesh/generate/twoD.h:139-179
void orc_mesh_cylinder(int nr, int nth, double rc, double rb, double* nodes)
{
    const double dr = (rb - rc) / (size_t)nr;
    const double dth = (2. * M_PI) / (size_t)nth;
    for (size_t i = 0; i <= (size_t)nr; ++i)
        for (size_t j = 0; j <= (size_t)nth; ++j) {
            const double r = rc + i * dr;
            const double th = j * dth;
            nodes[2 * (i * (nth + 1) + j)] = r * std::cos(th);
            nodes[2 * (i * (nth + 1) + j) + 1] = r * std::sin(th);
        }
}

/*
 * Geometric ratio s with dr0*(s^nr-1)/(s-1) = rb-rc (bisection to the last bit).
 * New generator next to make_cylinder_mesh (twoD.h:139-179): the O-grid of the
 * cylinder cases with geometric radial stretching.
 */
double orc_ogrid_ratio(int nr, double rc, double rb, double dr0)
{
    const double L = rb - rc;
    auto g = [&](double s) { return dr0 * (std::pow(s, nr) - 1.) / (s - 1.) - L; };
    double lo = 1. + 1e-12, hi = 2.;
    while (g(hi) < 0) hi *= 2.;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (g(mid) < 0) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

void orc_mesh_ogrid(int nr, int nth, double rc, double rb, double dr0, double* nodes)
{
    const double s = orc_ogrid_ratio(nr, rc, rb, dr0);
    for (int i = 0; i <= nr; ++i) {
        const double r = (i == nr) ? rb : rc + dr0 * (std::pow(s, i) - 1.) / (s - 1.);
        for (int j = 0; j <= nth; ++j) {
            const int jj = (j == nth) ? 0 : j;
            const double th = (2. * M_PI * jj) / nth;
            nodes[2 * ((size_t)i * (nth + 1) + j)] = r * std::cos(th);
            nodes[2 * ((size_t)i * (nth + 1) + j) + 1] = r * std::sin(th);
        }
    }
}

// conservationLaws/euler/species.ipp:2-18
orc_species orc_air_species(void)
{
    orc_species air;
    air.gamma = 1.4;
    air.minf = 0.0001;
    air.lref = 1.;
    air.pr = 0.7;
    air.nu = 1.81e-5;
    air.dt = std::numeric_limits<double>::max();
    air.R = 287.058;
    air.gamma1 = 1. / (air.gamma - 1.);
    return air;
}

void orc_set2state(const orc_case* cs, const double* q, double* state8)
{
    for (int k = 0; k < 8; ++k) state8[k] = 0.;
    set2state(*cs, to_species(cs->sp), q, state8);
}

void orc_state2set(const orc_case* cs, int basis, const double* state8, double* q)
{
    state2set(*cs, basis, state8, q);
}

/* delta transforms: 0 conserved, 1 primitive, 2 characteristic (Euler only) */
void orc_delta(const orc_case* cs, int from, int to, const double* st, const double* dq, double* out)
{
    const Species sp = to_species(cs->sp);
    double c[4];
    if (from == 0) for (int v = 0; v < 4; ++v) c[v] = dq[v];
    else if (from == 1) d_prim2cons(sp, st, dq, c);
    else { double p[4]; d_char2prim(sp, st, dq, p); d_prim2cons(sp, st, p, c); }
    if (to == 0) for (int v = 0; v < 4; ++v) out[v] = c[v];
    else if (to == 1) d_cons2prim(sp, st, c, out);
    else { double p[4]; d_cons2prim(sp, st, c, p); d_prim2char(sp, st, p, out); }
}

void orc_roe_average(const orc_case* cs, const double* sl8, const double* sr8, double* out8)
{
    for (int k = 0; k < 8; ++k) out8[k] = 0.;
    roe_average(to_species(cs->sp), sl8, sr8, out8);
}

static Face to_face(const orc_face* f)
{
    Face g;
    g.area = f->area;
    g.c[0] = f->cx;
    g.c[1] = f->cy;
    g.n[0] = f->nx;
    g.n[1] = f->ny;
    g.t[0] = f->tx;
    g.t[1] = f->ty;
    return g;
}

int orc_flux(const orc_case* cs, const orc_face* f, const double* ql, const double* qr, double* out)
{
    const FluxRes fr = flux_sets(*cs, to_species(cs->sp), to_face(f), ql, qr);
    const int nv = nvar_of(cs->law);
    for (int v = 0; v < nv; ++v) out[v] = fr.f[v];
    out[nv] = fr.lambda;
    return 0;
}

double orc_limit(int limiter, double a, double b) { return limit(limiter, a, b); }

/* which: 1 M1, 2 M2, 4 M4(param=beta), 11 P1, 13 P3, 15 P5(param=alpha) */
double orc_machsplit(int which, int sign, double m, double param)
{
    switch (which) {
    case 1: return msM1(sign, m);
    case 2: return msM2(sign, m);
    case 4: return msM4(sign, m, param);
    case 11: return msP1(sign, m);
    case 13: return msP3(sign, m);
    default: return msP5(sign, m, param);
    }
}

void orc_bc_update(const orc_case* cs, int bctype, const orc_face* f, const double* cell3,
                   const double* qint, const double* qref, double* out)
{
    Cell c;
    c.vol = cell3[0];
    c.c[0] = cell3[1];
    c.c[1] = cell3[2];
    bc_update(*cs, to_species(cs->sp), bctype, to_face(f), c, qint, qref, out);
}

void orc_wall_flux(const orc_case* cs, const orc_face* f, const double* qint, double* out)
{
    const FluxRes fr = wall_flux(*cs, to_species(cs->sp), to_face(f), qint);
    for (int v = 0; v < 4; ++v) out[v] = fr.f[v];
    out[4] = fr.lambda;
}

void orc_xmetrics(const orc_case* cs, int ni, int nj, const double* cells, double* xm)
{
    Mesh m;
    m.ni = ni;
    m.nj = nj;
    m.nodes = nullptr;
    m.cells.resize((size_t)ni * nj);
    for (size_t c = 0; c < (size_t)ni * nj; ++c) m.cells[c] = Cell{cells[3 * c], {cells[3 * c + 1], cells[3 * c + 2]}};
    std::vector<XM> X;
    xmetrics(*cs, m, X);
    for (size_t c = 0; c < X.size(); ++c)
        for (int k = 0; k < 3; ++k) xm[3 * c + k] = X[c].x[k];
}

void orc_qmetrics(const orc_case* cs, int ni, int nj, const double* cells, const double* q, double* qm)
{
    Mesh m;
    m.ni = ni;
    m.nj = nj;
    m.nodes = nullptr;
    m.cells.resize((size_t)ni * nj);
    for (size_t c = 0; c < (size_t)ni * nj; ++c) m.cells[c] = Cell{cells[3 * c], {cells[3 * c + 1], cells[3 * c + 2]}};
    std::vector<double> Q;
    qmetrics(*cs, m, q, Q);
    std::memcpy(qm, Q.data(), Q.size() * sizeof(double));
}

void orc_lsq_solve(const double* xm3, const double* qm, int nvar, const double* dir, double* out)
{
    lsq_solve(xm3, qm, nvar, dir, out);
}

void orc_muscl_face(const orc_case* cs, const orc_face* f, const double* cl3, const double* cr3,
                    const double* ql, const double* qr, const double* xl, const double* xr,
                    const double* qml, const double* qmr, double* out)
{
    const Cell cl{cl3[0], {cl3[1], cl3[2]}};
    const Cell cr{cr3[0], {cr3[1], cr3[2]}};
    const FluxRes fr = muscl_flux(*cs, to_species(cs->sp), to_face(f), cl, cr, ql, qr, xl, xr, qml, qmr);
    const int nv = nvar_of(cs->law);
    for (int v = 0; v < nv; ++v) out[v] = fr.f[v];
    out[nv] = fr.lambda;
}

int orc_residual(const orc_case* cs, int ni, int nj, const double* nodes, const double* q,
                 const double* bref, double* res)
{
    if (!valid_case(*cs, ni, nj)) return -1;
    Field F;
    setup_field(F, cs, ni, nj, nodes, bref);
    boundary_update(F, q);
    qmetrics(*cs, F.mesh, q, F.Q);
    std::vector<FluxRes> R;
    residual_calc(F, q, R);
    const int nv = F.nv;
    for (size_t c = 0; c < R.size(); ++c) {
        for (int v = 0; v < nv; ++v) res[c * (nv + 1) + v] = R[c].f[v];
        res[c * (nv + 1) + nv] = R[c].lambda;
    }
    return F.bc_ok ? 0 : -2;
}

double orc_spectral_radius(int ni, int nj, const double* cells, const double* res, int nvar)
{
    double r = std::numeric_limits<double>::min();
    for (size_t c = 0; c < (size_t)ni * nj; ++c) r = stdmax(r, res[c * (nvar + 1) + nvar] / cells[3 * c]);
    return r;
}

// timestepping/rungeKutta.h:37-167 and :172-197, spatial/eulerForwardUpdate.h:38-92
void orc_smooth_residual(const orc_case* cs, int ni, int nj, int mode, double alpha, double* res);

/*
 * integrate() (timestepping/rungeKutta.h:37-167); smooth_mode >= 0 smooths the accumulated
 * stage residual (orc_smooth_residual) before the forward-Euler update, the residual smoothing of
 * SteadyTimeControls (controls.h:31-48), which no integrate() of the reference consumes.
 */
static int integrate_impl(const orc_case* cs, int ni, int nj, const double* nodes, double* q0, const double* bref,
                          int nsteps, double* t_out, double* timers_ms, int smooth_mode, double smooth_alpha)
{
    if (!valid_case(*cs, ni, nj)) return -1;
    using clk = std::chrono::steady_clock;
    double tm[7] = {0, 0, 0, 0, 0, 0, 0};
    auto lap = [&](int k, clk::time_point& t0) {
        const auto t1 = clk::now();
        tm[k] += std::chrono::duration<double, std::milli>(t1 - t0).count();
        t0 = t1;
    };
    Field F;
    setup_field(F, cs, ni, nj, nodes, bref);
    const int nv = F.nv;
    const size_t N = (size_t)ni * nj;
    std::vector<double> q1(q0, q0 + N * nv), q2(q0, q0 + N * nv);
    std::vector<std::vector<FluxRes>> resStage(cs->nstages);
    std::vector<double> resTotal(N * nv);
    double t = 0;
    for (int tstep = 0; tstep < nsteps; ++tstep) {
        double dt{};
        for (int stg = 0; stg < cs->nstages; ++stg) {
            auto t0 = clk::now();
            boundary_update(F, q1.data());
            lap(0, t0);
            qmetrics(*cs, F.mesh, q1.data(), F.Q);
            lap(1, t0);
            residual_calc(F, q1.data(), resStage[stg]);
            lap(2, t0);
            if (stg == 0 && cs->dt_mode == ORC_DT_GLOBAL) dt = cs->cfl / spectral_radius_field(F.mesh, resStage[0]);
            if (stg == 0 && cs->dt_mode == ORC_DT_FIXED) dt = cs->cfl;
            lap(3, t0);
            // rungeKuttaAccumulation, rungeKutta.h:176-197
            const std::vector<FluxRes>* rs = resStage.data();
            const double* alpha = cs->alpha[stg];
#pragma omp parallel for schedule(static)
            for (long c = 0; c < (long)N; ++c) {
                double* rt = &resTotal[(size_t)c * nv];
                for (int v = 0; v < nv; ++v) rt[v] = 0.;
                for (int k = 0; k <= stg; ++k)
                    for (int v = 0; v < nv; ++v) rt[v] += rs[k][c].f[v] * alpha[k];
            }
            if (smooth_mode >= 0) {
                std::vector<double> r(N * (nv + 1));
                for (size_t c = 0; c < N; ++c) {
                    for (int v = 0; v < nv; ++v) r[c * (nv + 1) + v] = resTotal[c * nv + v];
                    r[c * (nv + 1) + nv] = rs[0][c].lambda;
                }
                orc_smooth_residual(cs, ni, nj, smooth_mode, smooth_alpha, r.data());
                for (size_t c = 0; c < N; ++c)
                    for (int v = 0; v < nv; ++v) resTotal[c * nv + v] = r[c * (nv + 1) + v];
            }
            lap(4, t0);
            const bool by_vol = cs->dt_mode != ORC_DT_LOCAL;   // global and fixed: (beta dt) / V
            const double stepfac = by_vol ? cs->beta[stg] * dt : cs->beta[stg] * cs->cfl;
#pragma omp parallel for schedule(static)
            for (long c = 0; c < (long)N; ++c) {
                const double fac = by_vol ? stepfac / F.mesh.cells[c].vol : stepfac / rs[0][c].lambda;
                double qc[MAXV];
                sol2cons(*cs, F.sp, &q0[(size_t)c * nv], qc);
                for (int v = 0; v < nv; ++v) qc[v] = qc[v] + resTotal[(size_t)c * nv + v] * fac;
                cons2sol(*cs, F.sp, qc, &q2[(size_t)c * nv]);
            }
            lap(5, t0);
            std::swap(q1, q2);
            lap(6, t0);
        }
        auto t0 = clk::now();
        std::memcpy(q0, q1.data(), N * nv * sizeof(double));
        lap(6, t0);
        t += dt;
    }
    if (t_out) *t_out = t;
    if (timers_ms) for (int k = 0; k < 7; ++k) timers_ms[k] = tm[k];
    return F.bc_ok ? 0 : -2;
}

int orc_integrate(const orc_case* cs, int ni, int nj, const double* nodes, double* q0,
                  const double* bref, int nsteps, double* t_out, double* timers_ms)
{
    return integrate_impl(cs, ni, nj, nodes, q0, bref, nsteps, t_out, timers_ms, -1, 1.0);
}

int orc_integrate_smoothed(const orc_case* cs, int ni, int nj, const double* nodes, double* q0, const double* bref,
                           int nsteps, int smooth_mode, double smooth_alpha, double* t_out)
{
    return integrate_impl(cs, ni, nj, nodes, q0, bref, nsteps, t_out, nullptr, smooth_mode, smooth_alpha);
}

void orc_norms(int ni, int nj, int nvar, const double* cells, const double* res, double* l2, double* linf)
{
    const size_t N = (size_t)ni * nj;
    for (int v = 0; v < nvar; ++v) {
        double s = 0., m = 0.;
        for (size_t c = 0; c < N; ++c) {
            const double r = res[c * (nvar + 1) + v] / cells[3 * c];
            s += r * r;
            m = std::fmax(m, std::fabs(r));
        }
        l2[v] = std::sqrt(s / (double)N);
        linf[v] = m;
    }
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* which: 11, 22, 33, 34 = ODE::Explicit::ssp11/22/33/34 (ode.h:107-197); 44 = classical RK4 */
void orc_rk_table(int which, int* nstages, double* alpha36, double* beta6)
{
    double a[6][6] = {};
    double b[6] = {};
    int ns = 0;
    if (which == 11) {
        ns = 1; b[0] = 1.0; a[0][0] = 1.0;
    } else if (which == 22) {
        ns = 2; b[0] = 1.0; b[1] = 1.0;
        a[0][0] = 1.0; a[1][0] = 0.5; a[1][1] = 0.5;
    } else if (which == 33) {
        ns = 3; b[0] = 1.0; b[1] = 0.5; b[2] = 1.0;
        a[0][0] = 1.; a[1][0] = 0.5; a[1][1] = 0.5;
        a[2][0] = 1. / 6.; a[2][1] = 1. / 6.; a[2][2] = 2. / 3.;
    } else if (which == 34) {
        ns = 4; b[0] = 0.5; b[1] = 1.0; b[2] = 0.5; b[3] = 1.0;
        a[0][0] = 1.0; a[1][0] = 0.5; a[1][1] = 0.5;
        a[2][0] = 1. / 3.; a[2][1] = 1. / 3.; a[2][2] = 1. / 3.;
        a[3][0] = 1. / 6.; a[3][1] = 1. / 6.; a[3][2] = 1. / 6.; a[3][3] = 1. / 2.;
    } else if (which == 44) {
        ns = 4; b[0] = 0.5; b[1] = 0.5; b[2] = 1.0; b[3] = 1.0;
        a[0][0] = 1.0; a[1][1] = 1.0; a[2][2] = 1.0;
        a[3][0] = 1. / 6.; a[3][1] = 1. / 3.; a[3][2] = 1. / 3.; a[3][3] = 1. / 6.;
    }
    *nstages = ns;
    for (int s = 0; s < 6; ++s) {
        beta6[s] = b[s];
        for (int k = 0; k < 6; ++k) alpha36[6 * s + k] = a[s][k];
    }
}

/*
 * Legacy gradient (spatial/gradientCalc.h:16-313): per cell and direction the sum of the
 * neighbour differences q_R - q_L over its edges (each edge added to both of its cells: i-edges
 * into dq[0], j-edges into dq[1], :153-203), then the periodic wraps (:247-313), whose loops stop
 * one short of the last row / column (j < nj-1, i < ni-1; quirk A-Q8).  dq [N][2][nvar].
 */
void orc_gradient(const orc_case* cs, int ni, int nj, const double* q, double* dq)
{
    const int nv = nvar_of(cs->law);
    const size_t N = (size_t)ni * nj;
    for (size_t k = 0; k < N * 2 * nv; ++k) dq[k] = 0.;
    auto add = [&](size_t cl, size_t cr, int dir) {
        for (int v = 0; v < nv; ++v) {
            const double d = q[cr * nv + v] - q[cl * nv + v];
            dq[(cl * 2 + dir) * nv + v] += d;
            dq[(cr * 2 + dir) * nv + v] += d;
        }
    };
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni - 1; ++i) add((size_t)i * nj + j, (size_t)(i + 1) * nj + j, 0);
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj - 1; ++j) add((size_t)i * nj + j, (size_t)i * nj + j + 1, 1);
    if (cs->bc[0] == ORC_BC_PERIODIC)
        for (int j = 0; j < nj - 1; ++j) add((size_t)(ni - 1) * nj + j, (size_t)j, 0);
    if (cs->bc[2] == ORC_BC_PERIODIC)
        for (int i = 0; i < ni - 1; ++i) add((size_t)i * nj + nj - 1, (size_t)i * nj, 1);
}

/*
 * Residual smoothing along one grid line of n values with stride `stride` (the reference's 1-D
 * operators, OLD.fieldOperations/centralExplicitSmoothing.ipp:5-39 and
 * centralImplicitSmoothing.ipp:2-60; beta = (alpha^2 - 1)/4, gamma = ((1 + 4 beta)/alpha - 1)/4):
 *   explicit: r1 = r0; for each edge (i-1, i): dr = gamma (r0[i] - r0[i-1]), r1[i-1] -= dr, r1[i] += dr,
 *             plus the wrap edge (n-1, 0) on a periodic line;
 *   implicit: 4 weighted (omega = 2/3) Jacobi sweeps of (1 + 2 beta) r - beta (r[i-1] + r[i+1]) = r0
 *             from r0 (the end values of a non-periodic line divide by 1 + beta), times alpha.
 * periodic: the line closes on itself ('p'), else the reference's 'a' ends.
 */
static void smooth_line(int mode, double alpha, bool periodic, int n, double* r, size_t stride)
{
    const double beta = 0.25 * (alpha * alpha - 1.);
    const double gamma = 0.25 * ((1. + 4. * beta) / alpha - 1.);
    std::vector<double> r0(n), r1(n), wrk(n);
    for (int i = 0; i < n; ++i) r0[i] = r[i * stride];
    if (mode == 0) {
        r1 = r0;
        for (int i = 1; i < n; ++i) {
            const double dr = gamma * (r0[i] - r0[i - 1]);
            r1[i - 1] -= dr;
            r1[i] += dr;
        }
        if (periodic) {
            const double dr = gamma * (r0[0] - r0[n - 1]);
            r1[n - 1] -= dr;
            r1[0] += dr;
        }
    } else {
        const double omega = 2. / 3.;
        wrk = r0;
        for (int it = 0; it < 4; ++it) {
            r1 = r0;
            for (int i = 0; i < n - 1; ++i) {
                r1[i] += beta * wrk[i + 1];
                r1[i + 1] += beta * wrk[i];
            }
            for (int i = 1; i < n - 1; ++i) r1[i] *= 1. / (1. + 2. * beta);
            if (periodic) {
                r1[0] += beta * wrk[n - 1];
                r1[n - 1] += beta * wrk[0];
                r1[0] *= 1. / (1. + 2. * beta);
                r1[n - 1] *= 1. / (1. + 2. * beta);
            } else {
                r1[0] *= 1. / (1. + 1. * beta);
                r1[n - 1] *= 1. / (1. + 1. * beta);
            }
            for (int i = 0; i < n; ++i) wrk[i] = (1. - omega) * wrk[i] + omega * r1[i];
        }
        r1 = wrk;
        for (int i = 0; i < n; ++i) r1[i] *= alpha;
    }
    for (int i = 0; i < n; ++i) r[i * stride] = r1[i];
}

/*
 * Residual smoothing of a residual field res [N][nvar+1] (the flux part; lambda is left alone),
 * applied as a product of the 1-D operators: along i on every j-line, then along j on every
 * i-line; a direction is periodic when its boundaries are.  mode 0 explicit, 1 implicit,
 * 2 explicit then implicit (the pairing of script/test/testSmooth.cpp:91-95).
 */
void orc_smooth_residual(const orc_case* cs, int ni, int nj, int mode, double alpha, double* res)
{
    const int nv = nvar_of(cs->law);
    const bool per_i = cs->bc[0] == ORC_BC_PERIODIC, per_j = cs->bc[2] == ORC_BC_PERIODIC;
    const int first = mode == 1 ? 1 : 0, last = mode == 0 ? 0 : 1;
    for (int m = first; m <= last; ++m)
        for (int v = 0; v < nv; ++v) {
            for (int j = 0; j < nj; ++j) smooth_line(m, alpha, per_i, ni, res + (size_t)j * (nv + 1) + v, (size_t)nj * (nv + 1));
            for (int i = 0; i < ni; ++i) smooth_line(m, alpha, per_j, nj, res + (size_t)i * nj * (nv + 1) + v, (size_t)(nv + 1));
        }
}

}   // extern "C"
