// hc_adjoint.hpp -- tangent (forward-mode) and adjoint (reverse-mode) residual kernels
// of the structured-mesh finite-volume path, FP64, one thread per face or per cell.
//
// The residual R(q) of one case is differentiated with respect to the solution q, the
// free-stream Mach number of the species (minf, used by AUSM+up and SLAU) and the
// boundary reference states (used by the far-field boundary conditions).
//
//   tangent:  given (q, qdot, minf_dot, bref_dot), the kernels return R(q) (bit-identical
//             to the stage kernel's residual: same operations in the same order) and
//             Rdot = dR/dq qdot + dR/dminf minf_dot + dR/dbref bref_dot, by evaluating the
//             point functions of hc_point.hpp on Dual<1> numbers.
//   adjoint:  given ybar, the kernels return qbar = (dR/dq)^T ybar, minf_bar and bref_bar.
//             Per face, the transposed flux Jacobian with respect to the reconstructed
//             left/right states (and minf, or the boundary reference state) is applied to
//             the face's adjoint weight: AUSM+up on interior faces by its hand-written
//             reverse sweep (hc_flux_vjp.hpp), the other fluxes from Dual<9> numbers, the
//             boundary faces one direction per lane from Dual<1> numbers in a kernel of
//             their own; the MUSCL reconstruction (limiter derivatives from Dual<2>, the
//             LSQ-biased slopes, which are linear in the LSQ solution metric Q) and the
//             metric Q = sum (q_R - q_L) dx are then transposed by hand.  A work-group
//             computes the faces of a tile of cells into LDS and its cells gather them in
//             a fixed order (deterministic, no atomics; AdjTile).
//
// Faces are numbered as the face-geometry arrays: i-faces l*nj + j (node line l, row j),
// then j-faces nfi + i*(nj+1) + l.  On a periodic direction the last node line is the
// first one and is skipped.  Semantics follow the stage kernel (hc_stage.hpp):
// spatial/muscl.h:10-87, spatial/residualCalc.h:199-305, ghostCellLoop.h:133-224,
// boundaryFluxLoop.h:150-261, spatial/lsqMetrics.h:181-213.
#pragma once

#include "hc_dual.hpp"
#include "hc_flux_vjp.hpp"
#include "hc_stage.hpp"

namespace hc {

// Runge-Kutta adjoint of one stage, given x = the adjoint of the stage's output state:
//   lam (+)= x  and  yb_k (+)= (coef_k / V_c) x  for the residuals k the stage combines,
// each target assigned on its first contribution of the step and added to afterwards
struct RkDeliver {
    double* lam;
    double* yb[6];
    double coef[6];
    int add[6];
    int nk;
    int lam_add;
};

HC_HD void rk_deliver(const RkDeliver& t, long long i, double x, double vol)
{
    t.lam[i] = t.lam_add ? t.lam[i] + x : x;
    for (int k = 0; k < t.nk; ++k) {
        const double y = (t.coef[k] / vol) * x;
        t.yb[k][i] = t.add[k] ? t.yb[k][i] + y : y;
    }
}

struct AdjArgs {
    MeshView<double> m;
    Species<double> sp;
    int vscale, pscale, limiter;
    int bc[4];
    long long N, nfaces;
    const double* q;       // [nv][N]
    const double* qd;      // tangent of q [nv][N]
    const double* bref;    // [rows][nv]
    const double* brefd;   // tangent of bref [rows][nv] or null
    double minfd;          // tangent of minf
    double* Q;             // LSQ solution metric [nv][2][N]
    double* Qd;            // its tangent [nv][2][N]
    double* F;             // face fluxes [nv+1][nfaces]
    double* Fd;            // their tangents [nv][nfaces]
    double* res;           // residual [N][nv+1] (layout of hc_residual)
    double* resd;          // its tangent [N][nv]
    const double* ybar;    // adjoint of the residual [nv][N]
    double* minf_face;     // per-face minf adjoint [nfaces], accumulated over calls (zeroed by the caller)
    double* brefbar;       // [rows][nv] (rows without a face stay as they are)
    double* qbar;          // [nv][N]
    double* Qbar;          // [nv][2][N]
    int bref_acc;          // 1: boundary faces add their reference-state adjoint to brefbar
    // adj_qedges delivers qbar: with rk_on as the Runge-Kutta adjoint rk of the previous stage,
    // else into qacc if set (qacc = qbar, or qacc = qacc + qbar with qacc_add), else into qbar
    int rk_on;
    RkDeliver rk;
    double* qacc;
    int qacc_add;
};

// geometry and cells of one face
struct FaceRef {
    int valid, isi, boundary, bid;
    long long fidx;   // index in the direction's face-slot arrays
    long long cl, cr; // left / right cells (boundary face: the interior cell in both)
    long long row;    // boundary reference row
};

HC_HD FaceRef face_ref(const MeshView<double>& m, long long f)
{
    FaceRef r{};
    const int ni = m.ni, nj = m.nj;
    if (f < m.nfi) {
        const int l = (int)(f / nj), j = (int)(f % nj);
        r.isi = 1;
        r.fidx = f;
        r.valid = !(m.per_i && l == ni);
        if ((l >= 1 && l <= ni - 1) || (m.per_i && l == 0)) {
            r.cl = (long long)((l - 1 + ni) % ni) * nj + j;
            r.cr = (long long)(l % ni) * nj + j;
        } else {
            r.boundary = 1;
            r.bid = l == 0 ? 0 : 1;
            r.cl = r.cr = (long long)(l == 0 ? 0 : ni - 1) * nj + j;
            r.row = (long long)(l == 0 ? 0 : nj) + j;
        }
    } else {
        const long long g = f - m.nfi;
        const int i = (int)(g / (nj + 1)), l = (int)(g % (nj + 1));
        r.isi = 0;
        r.fidx = g;
        r.valid = !(m.per_j && l == nj);
        if ((l >= 1 && l <= nj - 1) || (m.per_j && l == 0)) {
            r.cl = (long long)i * nj + (l - 1 + nj) % nj;
            r.cr = (long long)i * nj + l % nj;
        } else {
            r.boundary = 1;
            r.bid = l == 0 ? 2 : 3;
            r.cl = r.cr = (long long)i * nj + (l == 0 ? 0 : nj - 1);
            r.row = 2LL * nj + (l == 0 ? 0 : ni) + i;
        }
    }
    return r;
}

// faces of cell (i, j): left/right i-face, left/right j-face (periodic wrap onto node line 0)
HC_HD void cell_faces(const MeshView<double>& m, int i, int j, long long* fl)
{
    const int ni = m.ni, nj = m.nj;
    fl[0] = (long long)i * nj + j;
    fl[1] = (long long)((i + 1 == ni && m.per_i) ? 0 : i + 1) * nj + j;
    fl[2] = m.nfi + (long long)i * (nj + 1) + j;
    fl[3] = m.nfi + (long long)i * (nj + 1) + ((j + 1 == nj && m.per_j) ? 0 : j + 1);
}

template <class T> HC_HD T tval(double v, const double* d, long long idx)
{
    if constexpr (sizeof(T) == sizeof(double)) {
        return T(v);
    } else {
        T r(v);
        r.d[0] = d ? d[idx] : 0.0;
        return r;
    }
}

// LSQ solution metric of cell (gi, gj) in the stage kernel's order (hc_stage.hpp qmetric):
// interior i-edges, interior j-edges, periodic i wrap, periodic j wrap
template <class T, int NV>
HC_HD void qmetric_cell(const MeshView<double>& m, const double* q, const double* qd, int gi, int gj, T (*acc)[2])
{
    const int ni = m.ni, nj = m.nj;
    const long long N = (long long)ni * nj;
    for (int v = 0; v < NV; ++v) {
        acc[v][0] = T(0.0);
        acc[v][1] = T(0.0);
    }
    auto edge = [&](long long cl, long long cr, const double* base, long long nf, long long fidx) {
        const T dx0 = T(base[FS_DX0 * nf + fidx]);
        const T dx1 = T(base[FS_DX1 * nf + fidx]);
        for (int v = 0; v < NV; ++v) {
            const T dq = tval<T>(q[v * N + cr], qd, v * N + cr) - tval<T>(q[v * N + cl], qd, v * N + cl);
            acc[v][0] = acc[v][0] + dq * dx0;
            acc[v][1] = acc[v][1] + dq * dx1;
        }
    };
    const long long c = (long long)gi * nj + gj;
    if (gi > 0) edge(c - nj, c, m.fi, m.nfi, (long long)gi * nj + gj);
    if (gi < ni - 1) edge(c, c + nj, m.fi, m.nfi, (long long)(gi + 1) * nj + gj);
    if (gj > 0) edge(c - 1, c, m.fj, m.nfj, (long long)gi * (nj + 1) + gj);
    if (gj < nj - 1) edge(c, c + 1, m.fj, m.nfj, (long long)gi * (nj + 1) + gj + 1);
    if (m.per_i) {
        if (gi == 0) edge((long long)(ni - 1) * nj + gj, c, m.fi, m.nfi, (long long)gj);
        if (gi == ni - 1) edge(c, (long long)gj, m.fi, m.nfi, (long long)gj);
    }
    if (m.per_j) {
        if (gj == 0) edge((long long)gi * nj + nj - 1, c, m.fj, m.nfj, (long long)gi * (nj + 1));
        if (gj == nj - 1) edge(c, (long long)gi * nj, m.fj, m.nfj, (long long)gi * (nj + 1));
    }
}

// MUSCL face flux (spatial/muscl.h:10-87), the stage kernel's operations on type T
template <class T, int LAW, int BASIS, int FLUX>
HC_HD Flux4<T> muscl_flux(const Species<T>& sp, int vscale, int pscale, int limiter, const double* base, long long nf,
                          long long fidx, const T* ql, const T* qr, const T (*Ql)[2], const T (*Qr)[2])
{
    constexpr int NV = LawTraits<LAW>::nv;
    FaceG<T> f;
    f.area = T(base[FS_AREA * nf + fidx]);
    f.n0 = T(base[FS_N0 * nf + fidx]);
    f.n1 = T(base[FS_N1 * nf + fidx]);
    if (limiter == LIM_NOLIMIT1) return face_flux<T, LAW, BASIS, FLUX>(sp, vscale, pscale, f, ql, qr);
    const T dl = T(base[FS_DL * nf + fidx]);
    const T dr = T(base[FS_DR * nf + fidx]);
    const T s1 = T(base[FS_S1 * nf + fidx]);
    const T dx0 = T(base[FS_DX0 * nf + fidx]);
    const T dx1 = T(base[FS_DX1 * nf + fidx]);
    const T fl[3] = {T(base[FS_FL0 * nf + fidx]), T(base[FS_FL1 * nf + fidx]), T(base[FS_FL2 * nf + fidx])};
    const T fr[3] = {T(base[FS_FR0 * nf + fidx]), T(base[FS_FR1 * nf + fidx]), T(base[FS_FR2 * nf + fidx])};
    T qlf[NV], qrf[NV];
    for (int v = 0; v < NV; ++v) {
        const T dq = qr[v] - ql[v];
        const T dqdx = dq * s1;
        const T c0 = dq * dx0;
        const T c1 = dq * dx1;
        const T gl = lsq_solve1(fl, Ql[v][0] - c0, Ql[v][1] - c1, f.n0, f.n1);
        const T gr = lsq_solve1(fr, Qr[v][0] - c0, Qr[v][1] - c1, f.n0, f.n1);
        const T slope_l = limit(limiter, dqdx, gl);
        const T slope_r = limit(limiter, dqdx, gr);
        qlf[v] = ql[v] + slope_l * dl;
        qrf[v] = qr[v] - slope_r * dr;
    }
    return face_flux<T, LAW, BASIS, FLUX>(sp, vscale, pscale, f, qlf, qrf);
}

// non-periodic boundary face (the stage kernel's boundary_face on type T): ghost-cell flux
// or wall flux, oriented as the cell accumulation expects
template <class T, int LAW, int BASIS, int FLUX>
HC_HD Flux4<T> boundary_flux(const Species<T>& sp, int vscale, int pscale, int type, int bid, const double* base,
                             long long nf, long long fidx, const T* qint, const T* qref, double ccx, double ccy)
{
    constexpr int NV = LawTraits<LAW>::nv;
    FaceG<T> fres;
    fres.area = T(base[FS_AREA * nf + fidx]);
    fres.n0 = T(base[FS_N0 * nf + fidx]);
    fres.n1 = T(base[FS_N1 * nf + fidx]);
    FaceG<T> fupd = fres;
    if (bid == 1 || bid == 3) {
        fupd.n0 = -fres.n0;
        fupd.n1 = -fres.n1;
    }
    if constexpr (LAW == LAW_EULER) {
        if (type == BC_WALL) return euler_wall_flux<T, BASIS>(sp, fupd, qint);
    }
    T qg[NV];
    const T fcx = T(base[FS_CX * nf + fidx]);
    const T fcy = T(base[FS_CY * nf + fidx]);
    (void)ghost_value<T, LAW, BASIS>(sp, type, fupd, fcx, fcy, T(ccx), T(ccy), qint, qref, qg);
    if (bid == 0 || bid == 2) return face_flux<T, LAW, BASIS, FLUX>(sp, vscale, pscale, fres, qg, qint);
    return face_flux<T, LAW, BASIS, FLUX>(sp, vscale, pscale, fres, qint, qg);
}

HC_HD int bc_of(const AdjArgs& a, int bid)
{
    return bid == 0 ? a.bc[0] : bid == 1 ? a.bc[1] : bid == 2 ? a.bc[2] : a.bc[3];
}

template <int LAW, int BASIS, int FLUX> struct AdjKernels {
    static constexpr int NV = LawTraits<LAW>::nv;
    using D1 = Dual<1>;
    using D2 = Dual<2>;
    using D9 = Dual<9>;

    // ---- tangent: metric of q and of qdot ----
    HC_HD static void tan_qmetric(const AdjArgs& a, long long c)
    {
        const int gi = (int)(c / a.m.nj), gj = (int)(c % a.m.nj);
        D1 acc[NV][2];
        qmetric_cell<D1, NV>(a.m, a.q, a.qd, gi, gj, acc);
        for (int v = 0; v < NV; ++v)
            for (int d = 0; d < 2; ++d) {
                a.Q[(v * 2 + d) * a.N + c] = acc[v][d].v;
                a.Qd[(v * 2 + d) * a.N + c] = acc[v][d].d[0];
            }
    }

    // ---- tangent: flux of face f and its directional derivative ----
    HC_HD static void tan_face(const AdjArgs& a, long long f)
    {
        const FaceRef r = face_ref(a.m, f);
        if (!r.valid) return;
        Species<D1> sp = dual_species<1>(a.sp, -1);
        sp.minf.d[0] = a.minfd;
        const double* base = r.isi ? a.m.fi : a.m.fj;
        const long long nf = r.isi ? a.m.nfi : a.m.nfj;
        const long long N = a.N;
        Flux4<D1> F;
        if (!r.boundary) {
            D1 ql[NV], qr[NV], Ql[NV][2], Qr[NV][2];
            for (int v = 0; v < NV; ++v) {
                ql[v] = tval<D1>(a.q[v * N + r.cl], a.qd, v * N + r.cl);
                qr[v] = tval<D1>(a.q[v * N + r.cr], a.qd, v * N + r.cr);
                for (int d = 0; d < 2; ++d) {
                    Ql[v][d] = tval<D1>(a.Q[(v * 2 + d) * N + r.cl], a.Qd, (v * 2 + d) * N + r.cl);
                    Qr[v][d] = tval<D1>(a.Q[(v * 2 + d) * N + r.cr], a.Qd, (v * 2 + d) * N + r.cr);
                }
            }
            F = muscl_flux<D1, LAW, BASIS, FLUX>(sp, a.vscale, a.pscale, a.limiter, base, nf, r.fidx, ql, qr, Ql, Qr);
        } else {
            D1 qi[NV], qf[NV];
            for (int v = 0; v < NV; ++v) {
                qi[v] = tval<D1>(a.q[v * N + r.cl], a.qd, v * N + r.cl);
                qf[v] = tval<D1>(a.bref[r.row * NV + v], a.brefd, r.row * NV + v);
            }
            F = boundary_flux<D1, LAW, BASIS, FLUX>(sp, a.vscale, a.pscale, bc_of(a, r.bid), r.bid, base, nf, r.fidx, qi,
                                                     qf, a.m.ccx[r.cl], a.m.ccy[r.cl]);
        }
        for (int v = 0; v < NV; ++v) {
            a.F[v * a.nfaces + f] = F.f[v].v;
            a.Fd[v * a.nfaces + f] = F.f[v].d[0];
        }
        a.F[NV * a.nfaces + f] = F.lambda.v;
    }

    // ---- tangent: cell gather in the stage kernel's order (residualCalc.h:199-305) ----
    HC_HD static void tan_cell(const AdjArgs& a, long long c)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        const int i = (int)(c / nj), j = (int)(c % nj);
        long long fl[4];
        cell_faces(a.m, i, j, fl);
        double R[NV], Rd[NV], lam = 0.0;
        for (int v = 0; v < NV; ++v) R[v] = Rd[v] = 0.0;
        auto plus = [&](long long f) {
            for (int v = 0; v < NV; ++v) {
                R[v] = R[v] + a.F[v * a.nfaces + f];
                Rd[v] = Rd[v] + a.Fd[v * a.nfaces + f];
            }
            lam = lam + a.F[NV * a.nfaces + f];
        };
        auto minus = [&](long long f) {
            for (int v = 0; v < NV; ++v) {
                R[v] = R[v] - a.F[v * a.nfaces + f];
                Rd[v] = Rd[v] - a.Fd[v * a.nfaces + f];
            }
            lam = lam + a.F[NV * a.nfaces + f];
        };
        if (i > 0) plus(fl[0]);
        if (i < ni - 1) minus(fl[1]);
        if (j > 0) plus(fl[2]);
        if (j < nj - 1) minus(fl[3]);
        if (i == 0) plus(fl[0]);
        if (i == ni - 1) {
            if (a.bc[1] == BC_WALL) plus(fl[1]);
            else minus(fl[1]);
        }
        if (j == 0) plus(fl[2]);
        if (j == nj - 1) {
            if (a.bc[3] == BC_WALL) plus(fl[3]);
            else minus(fl[3]);
        }
        for (int v = 0; v < NV; ++v) {
            a.res[c * (NV + 1) + v] = R[v];
            a.resd[c * NV + v] = Rd[v];
        }
        a.res[c * (NV + 1) + NV] = lam;
    }

    // ---- adjoint: transpose of face f's flux and reconstruction ----
    // Contribution rows of a face: [0, NV) to the left cell, [NV, 2NV) to the right cell,
    // [2NV, 4NV) Qbar of the left cell (v*2+d), [4NV, 6NV) Qbar of the right cell.  A boundary face
    // gives its interior cell's contribution on the side that cell occupies (right of node line 0,
    // left of the last).

    // Non-periodic boundary face f, direction k of its transposed flux Jacobian: k < NV the interior
    // state, k < 2 NV the boundary reference state, k = 2 NV minf; one Dual<1> evaluation per
    // direction (each direction's derivative has the bits it has in a wider dual evaluation).
    // The directions of a face hand disjoint contribution rows to put(row, value): k < NV its
    // cell's row and a zero in the other side's row, NV <= k < 2 NV four of the (zero) Qbar rows;
    // they add the reference-state adjoint (k < 2 NV) to brefbar and the minf adjoint (k = 2 NV) to
    // minf_face[f].
    static constexpr int NDIR = 2 * NV + 1;
    template <class Put> HC_HD static void boundary_contrib(const AdjArgs& a, long long f, int k, Put&& put)
    {
        const FaceRef r = face_ref(a.m, f);
        const double* base = r.isi ? a.m.fi : a.m.fj;
        const long long nf = r.isi ? a.m.nfi : a.m.nfj;
        const long long N = a.N;
        const int type = bc_of(a, r.bid);
        const bool plus = (r.bid == 0 || r.bid == 2) || type == BC_WALL;
        const int side = (r.bid == 0 || r.bid == 2) ? NV : 0;   // interior cell right / left of the face
        if (k < NV) put(NV - side + k, 0.0);
        else if (k < 2 * NV)
            for (int m = 0; m < 4; ++m) put(2 * NV + 4 * (k - NV) + m, 0.0);
        const Species<D1> sp1 = dual_species<1>(a.sp, k == 2 * NV ? 0 : -1);
        D1 qi[NV], qf[NV];
        for (int v = 0; v < NV; ++v) {
            qi[v] = make_dual<1>(a.q[v * N + r.cl], v == k ? 0 : -1);
            qf[v] = make_dual<1>(a.bref[r.row * NV + v], NV + v == k ? 0 : -1);
        }
        const Flux4<D1> F = boundary_flux<D1, LAW, BASIS, FLUX>(sp1, a.vscale, a.pscale, type, r.bid, base, nf, r.fidx,
                                                                 qi, qf, a.m.ccx[r.cl], a.m.ccy[r.cl]);
        double g = 0.0;
        for (int v = 0; v < NV; ++v) {
            const double Fb = plus ? a.ybar[v * N + r.cl] : -a.ybar[v * N + r.cl];
            g = g + Fb * F.f[v].d[0];
        }
        if (k < NV) {
            put(side + k, g);
        } else if (k < 2 * NV) {
            double* dst = a.brefbar + r.row * NV + (k - NV);
            if (type != BC_WALL) *dst = a.bref_acc ? *dst + g : g;
        } else {
            a.minf_face[f] = a.minf_face[f] + g;
        }
    }

    // The 6 NV transposed contributions of a valid interior or periodic face, handed to put(row,
    // value) in the contribution rows, and the face's minf adjoint in *minf_out.  The left / right
    // cells' residual adjoints (yl, yr), states (ql, qr) and LSQ solution metrics (Ql, Qr) come in.
    template <class Put>
    HC_HD static void face_contrib(const AdjArgs& a, const FaceRef& r, const double* yl, const double* yr,
                                   const double* ql, const double* qr, const double (*Ql)[2], const double (*Qr)[2],
                                   double* minf_out, Put&& put)
    {
        const double* base = r.isi ? a.m.fi : a.m.fj;
        const long long nf = r.isi ? a.m.nfi : a.m.nfj;
        double Fb[NV];
        for (int v = 0; v < NV; ++v) Fb[v] = yr[v] - yl[v];
        const double area = base[FS_AREA * nf + r.fidx];
        const double n0 = base[FS_N0 * nf + r.fidx];
        const double n1 = base[FS_N1 * nf + r.fidx];
        double qlf[NV], qrf[NV];
        for (int v = 0; v < NV; ++v) {
            qlf[v] = ql[v];
            qrf[v] = qr[v];
        }
        // reconstruction (forward values and the limiter's partial derivatives)
        const bool recon = a.limiter != LIM_NOLIMIT1;
        double dl = 0, dr = 0, s1 = 0, dx0 = 0, dx1 = 0, fl[3] = {0, 0, 0}, fr[3] = {0, 0, 0};
        double Lla[NV], Llb[NV], Lra[NV], Lrb[NV];
        if (recon) {
            dl = base[FS_DL * nf + r.fidx];
            dr = base[FS_DR * nf + r.fidx];
            s1 = base[FS_S1 * nf + r.fidx];
            dx0 = base[FS_DX0 * nf + r.fidx];
            dx1 = base[FS_DX1 * nf + r.fidx];
            for (int k = 0; k < 3; ++k) {
                fl[k] = base[(FS_FL0 + k) * nf + r.fidx];
                fr[k] = base[(FS_FR0 + k) * nf + r.fidx];
            }
            for (int v = 0; v < NV; ++v) {
                const double Ql0 = Ql[v][0], Ql1 = Ql[v][1];
                const double Qr0 = Qr[v][0], Qr1 = Qr[v][1];
                const double dq = qr[v] - ql[v];
                const double dqdx = dq * s1;
                const double c0 = dq * dx0;
                const double c1 = dq * dx1;
                const double gl = lsq_solve1(fl, Ql0 - c0, Ql1 - c1, n0, n1);
                const double gr = lsq_solve1(fr, Qr0 - c0, Qr1 - c1, n0, n1);
                const double sl = limit_d(a.limiter, dqdx, gl, Lla[v], Llb[v]);
                const double sr = limit_d(a.limiter, dqdx, gr, Lra[v], Lrb[v]);
                qlf[v] = ql[v] + sl * dl;
                qrf[v] = qr[v] - sr * dr;
            }
        }
        // transposed flux Jacobian (reconstructed states and minf) applied to Fb
        double qlb[NV], qrb[NV], mb = 0.0;
        if constexpr (LAW == LAW_EULER && FLUX == FLUX_AUSMPUP) {
            for (int v = 0; v < NV; ++v) qlb[v] = qrb[v] = 0.0;
            mb = ausmpup_vjp<BASIS>(a.sp, a.vscale, a.pscale, FaceG<double>{area, n0, n1}, qlf, qrf, Fb, qlb, qrb);
        } else {
            const Species<D9> sp9 = dual_species<9>(a.sp, 8);
            FaceG<D9> fg;
            fg.area = D9(area);
            fg.n0 = D9(n0);
            fg.n1 = D9(n1);
            D9 xl[NV], xr[NV];
            for (int v = 0; v < NV; ++v) {
                xl[v] = make_dual<9>(qlf[v], v);
                xr[v] = make_dual<9>(qrf[v], 4 + v);
            }
            const Flux4<D9> F = face_flux<D9, LAW, BASIS, FLUX>(sp9, a.vscale, a.pscale, fg, xl, xr);
            for (int k = 0; k < NV; ++k) {
                double gl = 0.0, gr = 0.0;
                for (int v = 0; v < NV; ++v) {
                    gl = gl + Fb[v] * F.f[v].d[k];
                    gr = gr + Fb[v] * F.f[v].d[4 + k];
                }
                qlb[k] = gl;
                qrb[k] = gr;
            }
            for (int v = 0; v < NV; ++v) mb = mb + Fb[v] * F.f[v].d[8];
        }
        *minf_out = mb;
        if (!recon) {
            for (int v = 0; v < NV; ++v) {
                put(v, qlb[v]);
                put(NV + v, qrb[v]);
            }
            for (int k = 2 * NV; k < 6 * NV; ++k) put(k, 0.0);
            return;
        }
        // transpose of the reconstruction, variable by variable
        const double wl0 = fl[0] * n0 + fl[1] * n1, wl1 = fl[1] * n0 + fl[2] * n1;
        const double wr0 = fr[0] * n0 + fr[1] * n1, wr1 = fr[1] * n0 + fr[2] * n1;
        for (int v = 0; v < NV; ++v) {
            const double sbl = qlb[v] * dl;     // slope_l bar
            const double sbr = -qrb[v] * dr;    // slope_r bar
            const double dqdxb = sbl * Lla[v] + sbr * Lra[v];
            const double glb = sbl * Llb[v];
            const double grb = sbr * Lrb[v];
            const double u0l = glb * wl0, u1l = glb * wl1;
            const double u0r = grb * wr0, u1r = grb * wr1;
            const double c0b = -(u0l + u0r), c1b = -(u1l + u1r);
            const double dqb = dqdxb * s1 + c0b * dx0 + c1b * dx1;
            put(v, qlb[v] - dqb);
            put(NV + v, qrb[v] + dqb);
            put(2 * NV + 2 * v, u0l);
            put(2 * NV + 2 * v + 1, u1l);
            put(4 * NV + 2 * v, u0r);
            put(4 * NV + 2 * v + 1, u1r);
        }
    }

    // ---- adjoint: gather of the face contributions per cell (qbar direct part, Qbar) ----
    // cell c's direct adjoint and Qbar from its faces' contributions get(row, face) with face 0 / 1
    // the left / right i-face and 2 / 3 the left / right j-face: the cell is right of its left faces
    // and left of its right faces
    template <class Get> HC_HD static void gather(const AdjArgs& a, long long c, Get&& get)
    {
        for (int v = 0; v < NV; ++v) {
            double s = 0.0;
            s = s + get(NV + v, 0);
            s = s + get(v, 1);
            s = s + get(NV + v, 2);
            s = s + get(v, 3);
            a.qbar[v * a.N + c] = s;
            for (int d = 0; d < 2; ++d) {
                double t = 0.0;
                t = t + get(4 * NV + 2 * v + d, 0);
                t = t + get(2 * NV + 2 * v + d, 1);
                t = t + get(4 * NV + 2 * v + d, 2);
                t = t + get(2 * NV + 2 * v + d, 3);
                a.Qbar[(v * 2 + d) * a.N + c] = t;
            }
        }
    }

    // ---- adjoint: transpose of Q = sum over edges (q_R - q_L) dx ----
    HC_HD static void adj_qedges(const AdjArgs& a, long long c)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        const int gi = (int)(c / nj), gj = (int)(c % nj);
        const long long N = a.N;
        double add[NV];
        for (int v = 0; v < NV; ++v) add[v] = 0.0;
        // edge (A, B) with A left of B: the cell gains +s as B and -s as A
        auto edge = [&](long long A, long long B, const double* base, long long nf, long long fidx, double sign) {
            const double dx0 = base[FS_DX0 * nf + fidx], dx1 = base[FS_DX1 * nf + fidx];
            for (int v = 0; v < NV; ++v) {
                const double s = (a.Qbar[(v * 2) * N + A] + a.Qbar[(v * 2) * N + B]) * dx0 +
                                 (a.Qbar[(v * 2 + 1) * N + A] + a.Qbar[(v * 2 + 1) * N + B]) * dx1;
                add[v] = add[v] + sign * s;
            }
        };
        if (gi > 0) edge(c - nj, c, a.m.fi, a.m.nfi, (long long)gi * nj + gj, 1.0);
        else if (a.m.per_i) edge((long long)(ni - 1) * nj + gj, c, a.m.fi, a.m.nfi, (long long)gj, 1.0);
        if (gi < ni - 1) edge(c, c + nj, a.m.fi, a.m.nfi, (long long)(gi + 1) * nj + gj, -1.0);
        else if (a.m.per_i) edge(c, (long long)gj, a.m.fi, a.m.nfi, (long long)gj, -1.0);
        if (gj > 0) edge(c - 1, c, a.m.fj, a.m.nfj, (long long)gi * (nj + 1) + gj, 1.0);
        else if (a.m.per_j) edge((long long)gi * nj + nj - 1, c, a.m.fj, a.m.nfj, (long long)gi * (nj + 1), 1.0);
        if (gj < nj - 1) edge(c, c + 1, a.m.fj, a.m.nfj, (long long)gi * (nj + 1) + gj + 1, -1.0);
        else if (a.m.per_j) edge(c, (long long)gi * nj, a.m.fj, a.m.nfj, (long long)gi * (nj + 1), -1.0);
        for (int v = 0; v < NV; ++v) {
            const double qb = a.qbar[v * N + c] + add[v];
            if (a.rk_on) rk_deliver(a.rk, v * N + c, qb, a.m.cvol[c]);
            else if (!a.qacc) a.qbar[v * N + c] = qb;
            else a.qacc[v * N + c] = a.qacc_add ? a.qacc[v * N + c] + qb : qb;
        }
    }
};

// Adjoint faces and cell gather of one TI x TJ tile of cells in one work-group of NT threads, in
// the stage kernel's layout (hc_stage.hpp): phase 1 loads the state with a 2-cell halo and the
// residual adjoint with a 1-cell halo into LDS (periodic wrap); phase 2 forms the LSQ solution
// metric Q of the tile and its 1-cell ring in LDS (qmetric_cell's operations and order); phase 3
// computes the contributions of every interior and periodic face touching the tile into LDS
// (faces on a tile edge are computed by both tiles, with the same bits), then of the tile's
// non-periodic boundary faces, one lane per face and differentiation direction; phase 4 gathers
// each cell's direct adjoint and Qbar (AdjKernels::gather).  Neither Q nor the faces'
// contributions go through HBM.
template <int LAW, int BASIS, int FLUX, int TI, int TJ, int NT> struct AdjTile {
    using K = AdjKernels<LAW, BASIS, FLUX>;
    static constexpr int NV = LawTraits<LAW>::nv;
    static constexpr int QH = TI + 4, QW = TJ + 4;   // state tile with 2-cell halo
    static constexpr int MH = TI + 2, MW = TJ + 2;   // adjoint and metric tiles with 1-cell halo
    static constexpr int NFI = (TI + 1) * TJ;   // i-face of node line li, row lj: li * TJ + lj
    static constexpr int NFJ = TI * (TJ + 1);   // j-face of row li, node line lj: NFI + li * (TJ + 1) + lj
    static constexpr int NF = NFI + NFJ;
    struct Smem {
        double q[NV][QH][QW];
        double y[NV][MH][MW];
        double Q[NV][2][MH][MW];
        double c[6 * NV][NF];
    };

    // periodic wrap or clamp of a halo index (any distance: grids may be smaller than the halo)
    HC_HD static int wrap(int g, int n, int per)
    {
        if (per) return ((g % n) + n) % n;
        return g < 0 ? 0 : (g >= n ? n - 1 : g);
    }

    HC_HD static void load(const AdjArgs& a, Smem& s, int i0, int j0, int tid)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        const long long N = a.N;
        for (int k = tid; k < QH * QW; k += NT) {
            const int li = k / QW, lj = k % QW;
            const long long c = (long long)wrap(i0 - 2 + li, ni, a.m.per_i) * nj + wrap(j0 - 2 + lj, nj, a.m.per_j);
            for (int v = 0; v < NV; ++v) s.q[v][li][lj] = a.q[v * N + c];
            if (li >= 1 && li <= MH && lj >= 1 && lj <= MW)
                for (int v = 0; v < NV; ++v) s.y[v][li - 1][lj - 1] = a.ybar[v * N + c];
        }
    }

    // Q of the ring cell at metric position (mi, mj) = q-tile position (mi + 1, mj + 1): interior
    // i-edges, interior j-edges, periodic i wrap, periodic j wrap, as qmetric_cell
    HC_HD static void qmetric(const AdjArgs& a, Smem& s, int i0, int j0, int tid)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        for (int k = tid; k < MH * MW; k += NT) {
            const int mi = k / MW, mj = k % MW;
            const int ri = i0 - 1 + mi, rj = j0 - 1 + mj;
            if (!a.m.per_i && (ri < 0 || ri >= ni)) continue;
            if (!a.m.per_j && (rj < 0 || rj >= nj)) continue;
            const int gi = wrap(ri, ni, a.m.per_i), gj = wrap(rj, nj, a.m.per_j);
            const int r = mi + 1, c = mj + 1;
            double acc[NV][2];
            for (int v = 0; v < NV; ++v) acc[v][0] = acc[v][1] = 0.0;
            auto edge = [&](int rl, int cl, int rr, int cr, const double* base, long long nf, long long fidx) {
                const double dx0 = base[FS_DX0 * nf + fidx];
                const double dx1 = base[FS_DX1 * nf + fidx];
                for (int v = 0; v < NV; ++v) {
                    const double dq = s.q[v][rr][cr] - s.q[v][rl][cl];
                    acc[v][0] = acc[v][0] + dq * dx0;
                    acc[v][1] = acc[v][1] + dq * dx1;
                }
            };
            if (gi > 0) edge(r - 1, c, r, c, a.m.fi, a.m.nfi, (long long)gi * nj + gj);
            if (gi < ni - 1) edge(r, c, r + 1, c, a.m.fi, a.m.nfi, (long long)(gi + 1) * nj + gj);
            if (gj > 0) edge(r, c - 1, r, c, a.m.fj, a.m.nfj, (long long)gi * (nj + 1) + gj);
            if (gj < nj - 1) edge(r, c, r, c + 1, a.m.fj, a.m.nfj, (long long)gi * (nj + 1) + gj + 1);
            if (a.m.per_i) {
                if (gi == 0) edge(r - 1, c, r, c, a.m.fi, a.m.nfi, (long long)gj);
                if (gi == ni - 1) edge(r, c, r + 1, c, a.m.fi, a.m.nfi, (long long)gj);
            }
            if (a.m.per_j) {
                if (gj == 0) edge(r, c - 1, r, c, a.m.fj, a.m.nfj, (long long)gi * (nj + 1));
                if (gj == nj - 1) edge(r, c, r, c + 1, a.m.fj, a.m.nfj, (long long)gi * (nj + 1));
            }
            for (int v = 0; v < NV; ++v) {
                s.Q[v][0][mi][mj] = acc[v][0];
                s.Q[v][1][mi][mj] = acc[v][1];
            }
        }
    }

    HC_HD static void faces(const AdjArgs& a, Smem& s, int i0, int j0, int tid)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        for (int lf = tid; lf < NF; lf += NT) {
            const bool isi = lf < NFI;
            const int li = isi ? lf / TJ : (lf - NFI) / (TJ + 1);
            const int lj = isi ? lf % TJ : (lf - NFI) % (TJ + 1);
            const int gi = i0 + li, gj = j0 + lj;
            long long f;
            if (isi) {
                if (gi > ni || gj >= nj) continue;
                f = (long long)((gi == ni && a.m.per_i) ? 0 : gi) * nj + gj;
            } else {
                if (gi >= ni || gj > nj) continue;
                f = a.m.nfi + (long long)gi * (nj + 1) + ((gj == nj && a.m.per_j) ? 0 : gj);
            }
            const FaceRef r = face_ref(a.m, f);
            if (r.boundary) continue;
            // metric positions of the left and right cells (q-tile positions one further in)
            const int ml = isi ? li : li + 1, nl = isi ? lj + 1 : lj;
            const int mr = li + 1, nr = lj + 1;
            double yl[NV], yr[NV], ql[NV], qr[NV], Ql[NV][2], Qr[NV][2];
            for (int v = 0; v < NV; ++v) {
                yl[v] = s.y[v][ml][nl];
                yr[v] = s.y[v][mr][nr];
                ql[v] = s.q[v][ml + 1][nl + 1];
                qr[v] = s.q[v][mr + 1][nr + 1];
                for (int d = 0; d < 2; ++d) {
                    Ql[v][d] = s.Q[v][d][ml][nl];
                    Qr[v][d] = s.Q[v][d][mr][nr];
                }
            }
            double mb;
            K::face_contrib(a, r, yl, yr, ql, qr, Ql, Qr, &mb, [&](int k, double v) { s.c[k][lf] = v; });
            // one tile owns each face's minf adjoint: its first node line inside the tile (the
            // duplicate of a periodic wrap face at the far end is not the owner)
            if (isi ? (li < TI && gi < ni) : (lj < TJ && gj < nj)) a.minf_face[f] = a.minf_face[f] + mb;
        }
        // the tile's non-periodic boundary faces (each belongs to this tile alone), one lane per
        // face and direction: i-faces of node line 0 and ni, j-faces of node line 0 and nj
        const int lbi0 = (i0 == 0 && !a.m.per_i) ? 0 : -1;
        const int lbi1 = (!a.m.per_i && ni - i0 <= TI) ? ni - i0 : -1;
        const int lbj0 = (j0 == 0 && !a.m.per_j) ? 0 : -1;
        const int lbj1 = (!a.m.per_j && nj - j0 <= TJ) ? nj - j0 : -1;
        if (lbi0 < 0 && lbi1 < 0 && lbj0 < 0 && lbj1 < 0) return;
        for (int t = tid; t < 2 * (TI + TJ) * K::NDIR; t += NT) {
            const int b = t / K::NDIR, k = t % K::NDIR;
            int lf;
            long long f;
            if (b < 2 * TJ) {   // i-face of local node line lbi0 / lbi1, row lj
                const int li = b < TJ ? lbi0 : lbi1, lj = b % TJ;
                if (li < 0 || j0 + lj >= nj) continue;
                lf = li * TJ + lj;
                f = (long long)(i0 + li) * nj + j0 + lj;
            } else {            // j-face of row li, local node line lbj0 / lbj1
                const int bb = b - 2 * TJ;
                const int lj = bb < TI ? lbj0 : lbj1, li = bb % TI;
                if (lj < 0 || i0 + li >= ni) continue;
                lf = NFI + li * (TJ + 1) + lj;
                f = a.m.nfi + (long long)(i0 + li) * (nj + 1) + j0 + lj;
            }
            K::boundary_contrib(a, f, k, [&](int row, double v) { s.c[row][lf] = v; });
        }
    }

    HC_HD static void cells(const AdjArgs& a, const Smem& s, int i0, int j0, int tid)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        for (int lc = tid; lc < TI * TJ; lc += NT) {
            const int li = lc / TJ, lj = lc % TJ;
            const int gi = i0 + li, gj = j0 + lj;
            if (gi >= ni || gj >= nj) continue;
            const int lfl[4] = {li * TJ + lj, (li + 1) * TJ + lj, NFI + li * (TJ + 1) + lj, NFI + li * (TJ + 1) + lj + 1};
            K::gather(a, (long long)gi * nj + gj, [&](int row, int side) { return s.c[row][lfl[side]]; });
        }
    }
};

// tile shape of the fused adjoint kernel: 247 faces on 256 lanes, 70 KB of LDS in FP64 (two
// work-groups per CU, the occupancy its registers allow)
constexpr int ADJ_TI = 7, ADJ_TJ = 16, ADJ_NT = 256;

#if defined(__HIPCC__)

template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(ADJ_NT) k_adj_tile(AdjArgs a)
{
    using T = AdjTile<LAW, BASIS, FLUX, ADJ_TI, ADJ_TJ, ADJ_NT>;
    __shared__ typename T::Smem s;
    const int tiles_j = (a.m.nj + ADJ_TJ - 1) / ADJ_TJ;
    const int i0 = (int)(blockIdx.x / tiles_j) * ADJ_TI, j0 = (int)(blockIdx.x % tiles_j) * ADJ_TJ;
    T::load(a, s, i0, j0, threadIdx.x);
    __syncthreads();
    T::qmetric(a, s, i0, j0, threadIdx.x);
    __syncthreads();
    T::faces(a, s, i0, j0, threadIdx.x);
    __syncthreads();
    T::cells(a, s, i0, j0, threadIdx.x);
}

template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(256) k_tan_qmetric(AdjArgs a)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c < a.N) AdjKernels<LAW, BASIS, FLUX>::tan_qmetric(a, c);
}
template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(256) k_tan_face(AdjArgs a)
{
    const long long f = (long long)blockIdx.x * 256 + threadIdx.x;
    if (f < a.nfaces) AdjKernels<LAW, BASIS, FLUX>::tan_face(a, f);
}
template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(256) k_tan_cell(AdjArgs a)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c < a.N) AdjKernels<LAW, BASIS, FLUX>::tan_cell(a, c);
}
template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(256) k_adj_qedges(AdjArgs a)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c < a.N) AdjKernels<LAW, BASIS, FLUX>::adj_qedges(a, c);
}

#endif

}   // namespace hc
