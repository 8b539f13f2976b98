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
//             reverse sweep (hc_flux_vjp.hpp), the other fluxes and the boundary faces
//             from Dual<9> numbers (boundary faces in a kernel of their own, so that the
//             interior kernel does not hold the Dual<9> registers); the MUSCL reconstruction (limiter
//             derivatives from Dual<2>, the LSQ-biased slopes, which are linear in the
//             LSQ solution metric Q) and the metric Q = sum (q_R - q_L) dx are then
//             transposed by hand: faces write their contributions, cells gather them in
//             a fixed order (deterministic, no atomics).
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
    double* fout;          // per-face adjoint contributions [6 nv][nfaces]
    double* minf_face;     // per-face minf adjoint [nfaces]
    double* brefbar;       // [rows][nv] (rows without a face stay as they are)
    double* qbar;          // [nv][N]
    double* Qbar;          // [nv][2][N]
    int bref_acc;          // 1: boundary faces add their reference-state adjoint to brefbar
    double* qacc;          // if set, adj_qedges delivers qbar into it instead of qbar:
    int qacc_add;          //   qacc = qbar (0) or qacc = qacc + qbar (1)
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

    // ---- adjoint: metric of q (values only) ----
    HC_HD static void adj_qmetric(const AdjArgs& a, long long c)
    {
        const int gi = (int)(c / a.m.nj), gj = (int)(c % a.m.nj);
        double acc[NV][2];
        qmetric_cell<double, NV>(a.m, a.q, nullptr, gi, gj, acc);
        for (int v = 0; v < NV; ++v)
            for (int d = 0; d < 2; ++d) a.Q[(v * 2 + d) * a.N + c] = acc[v][d];
    }

    // ---- adjoint: transpose of face f's flux and reconstruction ----
    // fout rows: [0, NV) to the left cell, [NV, 2NV) to the right cell, [2NV, 4NV) Qbar of the
    // left cell (v*2+d), [4NV, 6NV) Qbar of the right cell.  A boundary face writes its interior
    // cell's contribution on the side that cell occupies (right of node line 0, left of the last).
    HC_HD static void adj_face(const AdjArgs& a, long long f)
    {
        if (face_ref(a.m, f).boundary)
            for (int k = 0; k < NDIR; ++k) adj_face_boundary(a, f, k);
        else
            adj_face_interior(a, f);
    }

    // boundary-face index b in [0, 2 nj + 2 ni) -> face index (i-faces of node lines 0 and ni, then
    // j-faces of node lines 0 and nj), whatever the directions' periodicity
    HC_HD static long long boundary_face(const MeshView<double>& m, long long b)
    {
        const long long ni = m.ni, nj = m.nj;
        if (b < nj) return b;
        if (b < 2 * nj) return ni * nj + (b - nj);
        if (b < 2 * nj + ni) return m.nfi + (b - 2 * nj) * (nj + 1);
        return m.nfi + (b - 2 * nj - ni) * (nj + 1) + nj;
    }

    // Non-periodic boundary face, direction k of its transposed flux Jacobian: k < NV the interior
    // state, k < 2 NV the boundary reference state, k = 2 NV minf; one Dual<1> evaluation per
    // direction (each direction's derivative has the bits it has in a wider dual evaluation).
    // The directions of a face write disjoint outputs: k < NV its cell's row and a zero in the
    // other side's row, NV <= k < 2 NV four of the Qbar rows (zeros) and the reference adjoint,
    // k = 2 NV the face's minf adjoint.  Other faces are left alone.
    static constexpr int NDIR = 2 * NV + 1;
    HC_HD static void adj_face_boundary(const AdjArgs& a, long long f, int k)
    {
        const FaceRef r = face_ref(a.m, f);
        if (!r.valid || !r.boundary) return;
        const long long nF = a.nfaces;
        const double* base = r.isi ? a.m.fi : a.m.fj;
        const long long nf = r.isi ? a.m.nfi : a.m.nfj;
        const long long N = a.N;
        const int type = bc_of(a, r.bid);
        const bool plus = (r.bid == 0 || r.bid == 2) || type == BC_WALL;
        const int side = (r.bid == 0 || r.bid == 2) ? NV : 0;   // interior cell right / left of the face
        if (k < NV) a.fout[(NV - side + k) * nF + f] = 0.0;
        else if (k < 2 * NV)
            for (int m = 0; m < 4; ++m) a.fout[(2 * NV + 4 * (k - NV) + m) * nF + f] = 0.0;
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
            a.fout[(side + k) * nF + f] = g;
        } else if (k < 2 * NV) {
            double* dst = a.brefbar + r.row * NV + (k - NV);
            if (type != BC_WALL) *dst = a.bref_acc ? *dst + g : g;
        } else {
            a.minf_face[f] = g;
        }
    }

    // interior or periodic face (boundary faces are left to adj_face_boundary)
    HC_HD static void adj_face_interior(const AdjArgs& a, long long f)
    {
        const FaceRef r = face_ref(a.m, f);
        if (r.boundary) return;
        const long long nF = a.nfaces;
        for (int k = 0; k < 6 * NV; ++k) a.fout[k * nF + f] = 0.0;
        a.minf_face[f] = 0.0;
        if (!r.valid) return;
        const double* base = r.isi ? a.m.fi : a.m.fj;
        const long long nf = r.isi ? a.m.nfi : a.m.nfj;
        const long long N = a.N;
        double Fb[NV];
        for (int v = 0; v < NV; ++v) Fb[v] = a.ybar[v * N + r.cr] - a.ybar[v * N + r.cl];
        const double area = base[FS_AREA * nf + r.fidx];
        const double n0 = base[FS_N0 * nf + r.fidx];
        const double n1 = base[FS_N1 * nf + r.fidx];
        double ql[NV], qr[NV], qlf[NV], qrf[NV];
        for (int v = 0; v < NV; ++v) {
            ql[v] = a.q[v * N + r.cl];
            qr[v] = a.q[v * N + r.cr];
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
                const double Ql0 = a.Q[(v * 2) * N + r.cl], Ql1 = a.Q[(v * 2 + 1) * N + r.cl];
                const double Qr0 = a.Q[(v * 2) * N + r.cr], Qr1 = a.Q[(v * 2 + 1) * N + r.cr];
                const double dq = qr[v] - ql[v];
                const double dqdx = dq * s1;
                const double c0 = dq * dx0;
                const double c1 = dq * dx1;
                const double gl = lsq_solve1(fl, Ql0 - c0, Ql1 - c1, n0, n1);
                const double gr = lsq_solve1(fr, Qr0 - c0, Qr1 - c1, n0, n1);
                const D2 sl = limit(a.limiter, make_dual<2>(dqdx, 0), make_dual<2>(gl, 1));
                const D2 sr = limit(a.limiter, make_dual<2>(dqdx, 0), make_dual<2>(gr, 1));
                Lla[v] = sl.d[0];
                Llb[v] = sl.d[1];
                Lra[v] = sr.d[0];
                Lrb[v] = sr.d[1];
                qlf[v] = ql[v] + sl.v * dl;
                qrf[v] = qr[v] - sr.v * dr;
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
        a.minf_face[f] = mb;
        if (!recon) {
            for (int v = 0; v < NV; ++v) {
                a.fout[v * nF + f] = qlb[v];
                a.fout[(NV + v) * nF + f] = qrb[v];
            }
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
            a.fout[v * nF + f] = qlb[v] - dqb;
            a.fout[(NV + v) * nF + f] = qrb[v] + dqb;
            a.fout[(2 * NV + 2 * v) * nF + f] = u0l;
            a.fout[(2 * NV + 2 * v + 1) * nF + f] = u1l;
            a.fout[(4 * NV + 2 * v) * nF + f] = u0r;
            a.fout[(4 * NV + 2 * v + 1) * nF + f] = u1r;
        }
    }

    // ---- adjoint: gather of the face contributions per cell (qbar direct part, Qbar) ----
    HC_HD static void adj_cell(const AdjArgs& a, long long c)
    {
        const int i = (int)(c / a.m.nj), j = (int)(c % a.m.nj);
        long long fl[4];
        cell_faces(a.m, i, j, fl);
        const long long nF = a.nfaces;
        // the cell is right of its left faces and left of its right faces
        for (int v = 0; v < NV; ++v) {
            double s = 0.0;
            s = s + a.fout[(NV + v) * nF + fl[0]];
            s = s + a.fout[v * nF + fl[1]];
            s = s + a.fout[(NV + v) * nF + fl[2]];
            s = s + a.fout[v * nF + fl[3]];
            a.qbar[v * a.N + c] = s;
            for (int d = 0; d < 2; ++d) {
                double t = 0.0;
                t = t + a.fout[(4 * NV + 2 * v + d) * nF + fl[0]];
                t = t + a.fout[(2 * NV + 2 * v + d) * nF + fl[1]];
                t = t + a.fout[(4 * NV + 2 * v + d) * nF + fl[2]];
                t = t + a.fout[(2 * NV + 2 * v + d) * nF + fl[3]];
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
            if (!a.qacc) a.qbar[v * N + c] = qb;
            else a.qacc[v * N + c] = a.qacc_add ? a.qacc[v * N + c] + qb : qb;
        }
    }
};

#if defined(__HIPCC__)

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
template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(256) k_adj_qmetric(AdjArgs a)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c < a.N) AdjKernels<LAW, BASIS, FLUX>::adj_qmetric(a, c);
}
template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(256) k_adj_face(AdjArgs a)
{
    const long long f = (long long)blockIdx.x * 256 + threadIdx.x;
    if (f < a.nfaces) AdjKernels<LAW, BASIS, FLUX>::adj_face_interior(a, f);
}
// one thread per direction of each of the 2 nj + 2 ni boundary-face positions
template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(64) k_adj_bface(AdjArgs a)
{
    using K = AdjKernels<LAW, BASIS, FLUX>;
    const long long t = (long long)blockIdx.x * 64 + threadIdx.x;
    const long long b = t / K::NDIR;
    if (b < 2LL * (a.m.ni + a.m.nj)) K::adj_face_boundary(a, K::boundary_face(a.m, b), (int)(t % K::NDIR));
}
template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(256) k_adj_cell(AdjArgs a)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c < a.N) AdjKernels<LAW, BASIS, FLUX>::adj_cell(a, c);
}
template <int LAW, int BASIS, int FLUX> __global__ void __launch_bounds__(256) k_adj_qedges(AdjArgs a)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c < a.N) AdjKernels<LAW, BASIS, FLUX>::adj_qedges(a, c);
}

#endif

}   // namespace hc
