// hc_stage.hpp -- the fused residual + Runge-Kutta stage of the structured-mesh
// finite-volume path, as one tiled kernel per stage.
//
// One work-group owns a TI x TJ tile of cells (j, the contiguous index, along
// the 64-lane wavefront).  Phases, separated by work-group barriers:
//   1. load   : solution variables of the tile plus a 2-cell halo into LDS
//               (periodic wrap across periodic boundaries)
//   2. qmetric: LSQ solution metric Q = sum dq (x) dx for the tile plus a
//               1-cell ring (spatial/lsqMetrics.h:181-213)
//   3. faces  : every i- and j-face touching the tile: MUSCL reconstruction
//               with LSQ-biased gradients and the limiter, then the flux
//               functor (spatial/muscl.h:10-87); boundary faces apply the ghost
//               value / wall flux of their boundary condition
//   4. cells  : owner-computes accumulation of the cell residual in the
//               reference's order (spatial/residualCalc.h:199-305), no atomics,
//               then the Runge-Kutta accumulation and forward-Euler update
//               (timestepping/rungeKutta.h:176-197, spatial/eulerForwardUpdate.h)
// The phases are __host__ __device__ so that a work-group can be executed
// thread by thread on the host for debugging; the kernel wrappers are at the end.
#pragma once

#include "hc_point.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

#include <cstdint>
#include <type_traits>
#include <cstring>

namespace hc {

// per-face geometry slots (slot-major arrays, one plane per slot)
enum FaceSlot {
    FS_AREA = 0, FS_N0, FS_N1, FS_DL, FS_DR, FS_S1, FS_DX0, FS_DX1,
    FS_FL0, FS_FL1, FS_FL2, FS_FR0, FS_FR1, FS_FR2, NSLOT
};
// boundary (non-periodic) faces store the face centre in FS_DL / FS_DR
constexpr int FS_CX = FS_DL;
constexpr int FS_CY = FS_DR;
// the FP32 slots of the mixed path hold the directional weights w = X'^-1 n of the biased LSQ
// factors in place of the factors (two per side)
constexpr int FS_WL0 = FS_FL0, FS_WL1 = FS_FL1, FS_WR0 = FS_FR0, FS_WR1 = FS_FR1;

// status bits returned through the C-ABI
enum StatusBits : unsigned int {
    ST_NONFINITE = 1u,     // non-finite state after an update
    ST_NONPHYSICAL = 2u,   // non-positive density or pressure (Euler)
    ST_BCNORMAL = 4u       // far-field boundary face normal does not point into the domain
};

template <class RF> struct MeshView {
    int ni, nj, per_i, per_j;
    long long nfi, nfj;   // number of i-faces (ni+1)*nj and j-faces ni*(nj+1)
    const double* cvol;
    const double* ccx;
    const double* ccy;
    const RF* fi;   // [NSLOT][nfi], i-face at node line l, row j: l*nj + j
    const RF* fj;   // [NSLOT][nfj], j-face at row i, node line l: i*(nj+1) + l
    const double* fi64;   // the same slots in FP64 (area and normal of the mixed path)
    const double* fj64;
    double period_i[2];   // translation from a cell to its periodic image across boundary 1
    double period_j[2];   // ... across boundary 3
};

// Runge-Kutta accumulation plan of one stage:
//   R_tot = (base >= 0 ? A[base] : +0) + R_s*alpha[s][s]
//   A[slot] = (first ? +0 : A[slot]) + R_s*alpha[r][s]   for later rows r
struct RkPlan {
    int base;
    int nupd;
    int upd_slot[5];
    int upd_first[5];
    double a_ss;
    double upd_coef[5];
    double beta;
};

enum StageMode { MODE_FUSED = 0, MODE_GLOBAL_STAGE0 = 1, MODE_RESIDUAL = 2 };

// A case's max(lambda/vol) of global-dt stage 0 is reduced into MAXSPREAD slots, each on its own
// 128-byte line (MAXSTRIDE words apart), one atomic per work-group on the slot of its tile index:
// atomics on one address serialise (a 4096^2 grid has 524288 waves), atomics on 64 lines do not.
// The update kernel takes the max of the slots with one load per lane of a wavefront.
constexpr int MAXSPREAD = 64, MAXSTRIDE = 16;

// Accumulation plans of a Runge-Kutta table in the reference's (alpha, beta) form
// (ode.h:87-105): one accumulator per later row with a non-zero coefficient
// before its diagonal.  Returns the number of accumulators.
template <class Vec> inline int make_rk_plans(int ns, const double (*alpha)[6], const double* beta, Vec& plans)
{
    int slot[6] = {-1, -1, -1, -1, -1, -1};
    int nslots = 0;
    for (int r = 1; r < ns; ++r) {
        bool any = false;
        for (int k = 0; k < r; ++k) any = any || alpha[r][k] != 0.0;
        if (any) slot[r] = nslots++;
    }
    plans.assign(ns, RkPlan{});
    for (int s = 0; s < ns; ++s) {
        RkPlan& p = plans[s];
        p.base = slot[s];
        p.a_ss = alpha[s][s];
        p.beta = beta[s];
        p.nupd = 0;
        for (int r = s + 1; r < ns; ++r) {
            if (slot[r] < 0 || alpha[r][s] == 0.0) continue;
            bool first = true;
            for (int k = 0; k < s; ++k) first = first && alpha[r][k] == 0.0;
            p.upd_slot[p.nupd] = slot[r];
            p.upd_first[p.nupd] = first ? 1 : 0;
            p.upd_coef[p.nupd] = alpha[r][s];
            p.nupd++;
        }
    }
    return nslots;
}

template <class RF> struct StageArgs {
    MeshView<RF> m;
    Species<double> sp;
    Species<RF> spf;
    int vscale, pscale, limiter;
    int bc[4];
    int mode, dt_mode, stage, reset_max;
    int keep_res;         // fused stages: also store the stage's residual in res_out (stage 0 of a step)
    long long N;
    double cfl;
    const double* q_in;   // stage state [nv][N]
    const double* q0;     // step-start state [nv][N]
    double* q_out;        // [nv][N]
    double* acc;          // accumulators [slot][nv][N]
    double* tmp;          // deferred R_tot of global-dt stage 0 [nv][N]
    double* lam0;         // stage-0 spectral radius sums [N]
    const double* bref;   // boundary reference values, AoS [2nj+2ni][nv]
    double* res_out;      // residual-only mode: AoS [N][nv+1]
    unsigned long long* maxslot;   // bits of max(lambda/vol), MAXSPREAD slots MAXSTRIDE apart, reset to bits(DBL_MIN)
    double* dtslot;       // [0] dt of the current step, [1] accumulated time
    unsigned int* status;
    RkPlan plan;
};

HC_HD unsigned long long dbl_bits(double x)
{
    unsigned long long u;
    std::memcpy(&u, &x, sizeof(u));
    return u;
}
HC_HD double bits_dbl(unsigned long long u)
{
    double x;
    std::memcpy(&x, &u, sizeof(x));
    return x;
}

// exponent-bit test: immune to floating-point contraction, which may turn x - x with x = a*b
// into fma(a, b, -a*b), the (non-zero) rounding error of the product
HC_HD bool hc_isfinite(double x) { return (dbl_bits(x) & 0x7FF0000000000000ULL) != 0x7FF0000000000000ULL; }

// Bijection of work-group ids onto tiles such that the ids an XCD receives
// (b % 8 == x) cover one contiguous band of tiles.
// the value of the first active lane, which the compiler then keeps in a scalar register: only
// for values that are equal on every active lane of the wavefront (identity on the host)
HC_HD int wave_uniform(int v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}

HC_HD int xcd_tile(int b, int ntiles)
{
    const int q = ntiles / 8, r = ntiles % 8;
    const int x = b % 8, y = b / 8;
    const int start = x * q + (x < r ? x : r);
    return start + y;
}

// NTHR threads per work-group, at least one per cell: the extra lanes share the load, metric and
// face phases, so more waves are resident per SIMD for the same LDS tile.
// UNI: uniform geometry (a periodic box whose face slots and cell volumes are all equal, checked
// bitwise on the host): every face takes the slots of face 0 of its kind and every cell the volume
// of cell 0, read at wave-uniform addresses (scalar loads) instead of one load per face and slot.
template <int LAW, int BASIS, int FLUX, class RF, int TI, int TJ, int NTHR = TI * TJ, int CST = 0, int GPF = 0,
          int UNI = 0>
struct StageTile {
    static constexpr int NV = LawTraits<LAW>::nv;
    static constexpr int NC = TI * TJ;   // cells of the tile
    static constexpr int NT = NTHR;      // threads of the work-group
    static_assert(NT >= NC, "one thread per cell at least");
    static constexpr int QH = TI + 4, QW = TJ + 4;   // q tile with 2-cell halo
    static constexpr int MH = TI + 2, MW = TJ + 2;   // Q ring with 1-cell halo
    static constexpr int NFI = (TI + 1) * TJ;
    static constexpr int NFJ = TI * (TJ + 1);
    static constexpr int NF = NFI + NFJ;

    // Mixed precision (RF = float) with a splitting flux (mixed_wide): the solution tile, the face
    // states, area and normal, the central flux and the stored face fluxes stay FP64 (O(1)
    // quantities); the LSQ metric, the slopes, the limiter and the flux dissipation, which scale
    // with the jumps, run in FP32.  Otherwise RF = float throughout the faces.
    static constexpr bool MIXED = sizeof(RF) == sizeof(float) && mixed_wide(LAW, FLUX);
    using QT = std::conditional_t<MIXED, double, RF>;   // solution tile
    using FT = std::conditional_t<MIXED, double, RF>;   // face fluxes

    // Centre staging (FP64 and wide mixed kernels): the cell centres of the q tile sit in LDS and
    // the centre differences dx_c of the metric edges and faces are formed there, with the
    // host geometry's operations (c_R - c_L, and c_R - (c_L - period) across a periodic wrap),
    // instead of being read as two face slots per edge.
    static constexpr bool CSTAGE = CST != 0 && UNI == 0 && (sizeof(RF) == sizeof(double) || MIXED);
    static constexpr int CH = CSTAGE ? QH : 1, CW = CSTAGE ? QW : 1;

    struct Smem {
        QT q[NV][QH][QW];
        RF Q[NV][2][MH][MW];
        FT F[NV + 1][NF];       // face fluxes and spectral radii, i-faces then j-faces
        double gc[2][CH][CW];   // cell centres of the q tile
    };

    // periodic wrap or clamp of a halo index; on grids larger than the halo reach the wrap is
    // one step of n, and the integer division runs only on lanes of smaller grids
    HC_HD static int wrapi(int g, int n, int per)
    {
        if (per) {
            if (g < 0) g += n;
            else if (g >= n) g -= n;
            if ((unsigned)g >= (unsigned)n) g = ((g % n) + n) % n;
            return g;
        }
        return g < 0 ? 0 : (g >= n ? n - 1 : g);
    }

    // ---- phase 1: solution tile with a 2-cell halo ----
    HC_HD static void load(const StageArgs<RF>& a, Smem& s, int i0, int j0, int tid)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        for (int k = tid; k < QH * QW; k += NT) {
            const int li = k / QW, lj = k % QW;
            const int gi = wrapi(i0 - 2 + li, ni, a.m.per_i);
            const int gj = wrapi(j0 - 2 + lj, nj, a.m.per_j);
            const long long c = (long long)gi * nj + gj;
            for (int v = 0; v < NV; ++v) s.q[v][li][lj] = QT(a.q_in[v * a.N + c]);
            if constexpr (CSTAGE) {
                s.gc[0][li][lj] = a.m.ccx[c];
                s.gc[1][li][lj] = a.m.ccy[c];
            }
        }
    }

    // centre difference of the cells at q-tile positions pl, pr: c_R - c_L, or across a periodic
    // wrap c_R - (c_L - period) (the host geometry's wrap-face operations)
    HC_HD static void centre_diff(const StageArgs<RF>& a, const Smem& s, int pl, int pr, bool wrap, bool isi,
                                  double& d0, double& d1)
    {
        const double* gcx = &s.gc[0][0][0];
        const double* gcy = &s.gc[1][0][0];
        double lx = gcx[pl], ly = gcy[pl];
        if (wrap) {
            const double* per = isi ? a.m.period_i : a.m.period_j;
            lx = lx - per[0];
            ly = ly - per[1];
        }
        d0 = gcx[pr] - lx;
        d1 = gcy[pr] - ly;
    }

    // ---- phase 2: LSQ solution metric of tile cells and the 1-cell ring ----
    // per cell: interior i-edges, interior j-edges, periodic i wrap, periodic j wrap
    // (the oracle's neighbour_accumulation order, parallalg/neighbour_algorithm.h:253-317)
    HC_HD static void qmetric(const StageArgs<RF>& a, Smem& s, int i0, int j0, int tid)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        const RF* fi = a.m.fi;
        const RF* fj = a.m.fj;
        const long long nfi = a.m.nfi, nfj = a.m.nfj;
        for (int k = tid; k < MH * MW; k += NT) {
            const int li = k / MW, lj = k % MW;
            const int ri = i0 - 1 + li, rj = j0 - 1 + lj;   // unwrapped global index
            if (!a.m.per_i && (ri < 0 || ri >= ni)) continue;
            if (!a.m.per_j && (rj < 0 || rj >= nj)) continue;
            const int gi = wrapi(ri, ni, a.m.per_i);
            const int gj = wrapi(rj, nj, a.m.per_j);
            const int r = li + 1, c = lj + 1;   // position in the q tile
            RF acc[NV][2];
            for (int v = 0; v < NV; ++v) {
                acc[v][0] = RF(0);
                acc[v][1] = RF(0);
            }
            auto edge = [&](int rl, int cl, int rr, int cr, const RF* base, long long nf, long long fidx, bool wrap) {
                RF dx0, dx1;
                if constexpr (UNI != 0) {
                    dx0 = base[FS_DX0 * nf];
                    dx1 = base[FS_DX1 * nf];
                } else if constexpr (CSTAGE) {
                    double d0, d1;
                    centre_diff(a, s, rl * QW + cl, rr * QW + cr, wrap, cl == cr, d0, d1);
                    dx0 = RF(d0);
                    dx1 = RF(d1);
                } else {
                    dx0 = base[FS_DX0 * nf + fidx];
                    dx1 = base[FS_DX1 * nf + fidx];
                }
                for (int v = 0; v < NV; ++v) {
                    const RF dq = RF(s.q[v][rr][cr] - s.q[v][rl][cl]);
                    acc[v][0] = acc[v][0] + dq * dx0;
                    acc[v][1] = acc[v][1] + dq * dx1;
                }
            };
            if (gi > 0) edge(r - 1, c, r, c, fi, nfi, (long long)gi * nj + gj, false);
            if (gi < ni - 1) edge(r, c, r + 1, c, fi, nfi, (long long)(gi + 1) * nj + gj, false);
            if (gj > 0) edge(r, c - 1, r, c, fj, nfj, (long long)gi * (nj + 1) + gj, false);
            if (gj < nj - 1) edge(r, c, r, c + 1, fj, nfj, (long long)gi * (nj + 1) + gj + 1, false);
            if (a.m.per_i) {
                if (gi == 0) edge(r - 1, c, r, c, fi, nfi, (long long)gj, true);
                if (gi == ni - 1) edge(r, c, r + 1, c, fi, nfi, (long long)gj, true);
            }
            if (a.m.per_j) {
                if (gj == 0) edge(r, c - 1, r, c, fj, nfj, (long long)gi * (nj + 1), true);
                if (gj == nj - 1) edge(r, c, r, c + 1, fj, nfj, (long long)gi * (nj + 1), true);
            }
            for (int v = 0; v < NV; ++v) {
                s.Q[v][0][li][lj] = acc[v][0];
                s.Q[v][1][li][lj] = acc[v][1];
            }
        }
    }

    // Geometry of one interior or periodic face: its slots (area and normal in FP64 for the
    // mixed path) and the centre difference
    using GT = std::conditional_t<MIXED, double, RF>;
    struct Geo {
        GT area, n0, n1;
        RF dl, dr, s1, dx0, dx1, fl[3], fr[3];
    };

    // (cx0, cx1): the centre difference formed in LDS (centre staging), else read from the slots
    HC_HD static Geo load_geo(const RF* base, const double* base64, long long nf, long long fidx, double cx0,
                              double cx1)
    {
        Geo g;
        if constexpr (MIXED) {
            g.area = base64[FS_AREA * nf + fidx];
            g.n0 = base64[FS_N0 * nf + fidx];
            g.n1 = base64[FS_N1 * nf + fidx];
        } else {
            g.area = base[FS_AREA * nf + fidx];
            g.n0 = base[FS_N0 * nf + fidx];
            g.n1 = base[FS_N1 * nf + fidx];
        }
        g.dl = base[FS_DL * nf + fidx];
        g.dr = base[FS_DR * nf + fidx];
        g.s1 = base[FS_S1 * nf + fidx];
        g.dx0 = CSTAGE ? RF(cx0) : base[FS_DX0 * nf + fidx];
        g.dx1 = CSTAGE ? RF(cx1) : base[FS_DX1 * nf + fidx];
        if constexpr (sizeof(RF) == sizeof(float)) {
            g.fl[0] = base[FS_WL0 * nf + fidx];
            g.fl[1] = base[FS_WL1 * nf + fidx];
            g.fr[0] = base[FS_WR0 * nf + fidx];
            g.fr[1] = base[FS_WR1 * nf + fidx];
        } else {
            for (int d = 0; d < 3; ++d) {
                g.fl[d] = base[(FS_FL0 + d) * nf + fidx];
                g.fr[d] = base[(FS_FR0 + d) * nf + fidx];
            }
        }
        return g;
    }

    // uniform geometry: the slots of i-face 0 and j-face 0, each read at one wave-uniform address,
    // and the lane's kind selected
    HC_HD static Geo uni_geo(const StageArgs<RF>& a, bool isi)
    {
        const Geo gi = load_geo(a.m.fi, a.m.fi64, a.m.nfi, 0, 0.0, 0.0);
        const Geo gj = load_geo(a.m.fj, a.m.fj64, a.m.nfj, 0, 0.0, 0.0);
        Geo g;
        g.area = isi ? gi.area : gj.area;
        g.n0 = isi ? gi.n0 : gj.n0;
        g.n1 = isi ? gi.n1 : gj.n1;
        g.dl = isi ? gi.dl : gj.dl;
        g.dr = isi ? gi.dr : gj.dr;
        g.s1 = isi ? gi.s1 : gj.s1;
        g.dx0 = isi ? gi.dx0 : gj.dx0;
        g.dx1 = isi ? gi.dx1 : gj.dx1;
        for (int d = 0; d < 3; ++d) {
            g.fl[d] = isi ? gi.fl[d] : gj.fl[d];
            g.fr[d] = isi ? gi.fr[d] : gj.fr[d];
        }
        return g;
    }

    // gradient along n from the biased LSQ metric: the factor solve on FP64 slots
    // (lsq/factor_solve.ipp:47-69), the directional weights on the FP32 slots
    HC_HD static RF directional(const RF* f, RF q0, RF q1, RF n0, RF n1)
    {
        if constexpr (sizeof(RF) == sizeof(float)) return q0 * f[0] + q1 * f[1];
        else return lsq_solve1(f, q0, q1, n0, n1);
    }

    // MUSCL face flux (spatial/muscl.h:10-87) with precomputed face geometry
    HC_HD static Flux4<FT> muscl(const StageArgs<RF>& a, const Geo& g, const QT* ql, const QT* qr,
                                 const RF (*Ql)[2], const RF (*Qr)[2])
    {
        if constexpr (MIXED) {
            const FaceG<double> f{g.area, g.n0, g.n1};
            if (a.limiter == LIM_NOLIMIT1)
                return face_flux_mixed<LAW, BASIS, FLUX>(a.sp, a.spf, a.vscale, a.pscale, f, ql, qr);
            const RF n0 = RF(f.n0), n1 = RF(f.n1);
            RF dqdx[NV], gl[NV], gr[NV], sl[NV], sr[NV];
            for (int v = 0; v < NV; ++v) {
                const RF dq = RF(qr[v] - ql[v]);
                dqdx[v] = dq * g.s1;
                const RF c0 = dq * g.dx0;
                const RF c1 = dq * g.dx1;
                gl[v] = directional(g.fl, Ql[v][0] - c0, Ql[v][1] - c1, n0, n1);
                gr[v] = directional(g.fr, Qr[v][0] - c0, Qr[v][1] - c1, n0, n1);
            }
            limit_pairs<RF, NV>(a.limiter, dqdx, gl, gr, sl, sr);
            double qlf[NV], qrf[NV];
            for (int v = 0; v < NV; ++v) {
                qlf[v] = ql[v] + double(sl[v] * g.dl);
                qrf[v] = qr[v] - double(sr[v] * g.dr);
            }
            return face_flux_mixed<LAW, BASIS, FLUX>(a.sp, a.spf, a.vscale, a.pscale, f, qlf, qrf);
        } else {
            const FaceG<RF> f{g.area, g.n0, g.n1};
            if (a.limiter == LIM_NOLIMIT1) return face_flux<RF, LAW, BASIS, FLUX>(a.spf, a.vscale, a.pscale, f, ql, qr);
            RF dqdx[NV], gl[NV], gr[NV], sl[NV], sr[NV];
            for (int v = 0; v < NV; ++v) {
                const RF dq = qr[v] - ql[v];
                dqdx[v] = dq * g.s1;
                const RF c0 = dq * g.dx0;
                const RF c1 = dq * g.dx1;
                gl[v] = directional(g.fl, Ql[v][0] - c0, Ql[v][1] - c1, f.n0, f.n1);
                gr[v] = directional(g.fr, Qr[v][0] - c0, Qr[v][1] - c1, f.n0, f.n1);
            }
            // one branch on the limiter per face, straight-line code over the variables
            limit_pairs<RF, NV>(a.limiter, dqdx, gl, gr, sl, sr);
            RF qlf[NV], qrf[NV];
            for (int v = 0; v < NV; ++v) {
                qlf[v] = ql[v] + sl[v] * g.dl;
                qrf[v] = qr[v] - sr[v] * g.dr;
            }
            return face_flux<RF, LAW, BASIS, FLUX>(a.spf, a.vscale, a.pscale, f, qlf, qrf);
        }
    }

    // non-periodic boundary face: ghost-cell flux or wall flux, in the orientation the
    // cell accumulation expects (ghostCellLoop.h:133-224, boundaryFluxLoop.h:150-261); the mixed
    // path evaluates its (few) boundary faces in FP64
    HC_HD static Flux4<FT> boundary_face(const StageArgs<RF>& a, int bid, const RF* base, const double* base64,
                                         long long nf, long long fidx, const QT* qint, long long cint,
                                         long long bref_row, unsigned int& flag)
    {
        using BT = std::conditional_t<MIXED, double, RF>;
        const BT* gb;
        if constexpr (MIXED) gb = base64;
        else gb = base;
        const Species<BT>* sp;
        if constexpr (MIXED) sp = &a.sp;
        else sp = &a.spf;
        FaceG<BT> fres;
        fres.area = gb[FS_AREA * nf + fidx];
        fres.n0 = gb[FS_N0 * nf + fidx];
        fres.n1 = gb[FS_N1 * nf + fidx];
        // the update surface faces into the domain: reversed on boundaries 1 and 3
        FaceG<BT> fupd = fres;
        if (bid == 1 || bid == 3) {
            fupd.n0 = -fres.n0;
            fupd.n1 = -fres.n1;
        }
        // constant indices only: the arguments may live in registers
        const int type = bid == 0 ? a.bc[0] : bid == 1 ? a.bc[1] : bid == 2 ? a.bc[2] : a.bc[3];
        if constexpr (LAW == LAW_EULER) {
            if (type == BC_WALL) return euler_wall_flux<BT, BASIS>(*sp, fupd, qint);
        }
        BT qref[NV], qg[NV];
        for (int v = 0; v < NV; ++v) qref[v] = BT(a.bref[bref_row * NV + v]);
        const BT fcx = gb[FS_CX * nf + fidx];
        const BT fcy = gb[FS_CY * nf + fidx];
        const BT ccx = BT(a.m.ccx[cint]);
        const BT ccy = BT(a.m.ccy[cint]);
        if (!ghost_value<BT, LAW, BASIS>(*sp, type, fupd, fcx, fcy, ccx, ccy, qint, qref, qg)) flag |= ST_BCNORMAL;
        if (bid == 0 || bid == 2) return face_flux<BT, LAW, BASIS, FLUX>(*sp, a.vscale, a.pscale, fres, qg, qint);
        return face_flux<BT, LAW, BASIS, FLUX>(*sp, a.vscale, a.pscale, fres, qint, qg);
    }

    HC_HD static void store(Smem& s, int k, const Flux4<FT>& F)
    {
        for (int v = 0; v < NV; ++v) s.F[v][k] = F.f[v];
        s.F[NV][k] = F.lambda;
    }

    // ---- phase 3: all faces of the tile ----
    // One face index space: i-faces k = li*TJ + lj (node line i0+li, row j0+lj), then j-faces
    // k = NFI + li*(TJ+1) + lj (row i0+li, node line j0+lj).  Both kinds run the same MUSCL and
    // flux code on lane-selected operands, so a wave holding both does not diverge.
    // Geometry prefetch: the slots of the lane's first face (k = tid), loaded while the tile and
    // the metric phase are in flight (their latency otherwise opens the face phase); the centre
    // difference of centre staging is filled in the face phase
    // Face k of the tile: its kind and its position (li, lj) in the kind's layout.  When the i-faces
    // fill whole wavefronts (NFI a multiple of 64), the 64 consecutive k of a wavefront are of one
    // kind, so the kind is wave-uniform: the operands it selects (face-slot arrays, strides,
    // periods, LDS offsets) are then picked with scalar selects instead of per-lane ones.
    static constexpr bool KIND_UNIFORM = NFI % 64 == 0 && NT % 64 == 0;
    HC_HD static bool face_kind_i(int k)
    {
        if constexpr (KIND_UNIFORM) return wave_uniform(k) < NFI;
        else return k < NFI;
    }
    HC_HD static void face_pos(int k, bool isi, int& li, int& lj)
    {
        if (isi) {
            li = k / TJ;
            lj = k % TJ;
        } else {
            li = (k - NFI) / (TJ + 1);
            lj = (k - NFI) % (TJ + 1);
        }
    }

    HC_HD static void prefetch(const StageArgs<RF>& a, int i0, int j0, int tid, Geo& g)
    {
        if constexpr (GPF != 0 && UNI == 0) {
            const int k = tid;
            if (k >= NF) return;
            const bool isi = face_kind_i(k);
            int li, lj;
            face_pos(k, isi, li, lj);
            const int l = isi ? i0 + li : j0 + lj;
            const int t = isi ? j0 + lj : i0 + li;
            const int lmax = isi ? a.m.ni : a.m.nj;
            if (l > lmax || t >= (isi ? a.m.nj : a.m.ni)) return;
            const bool interior = l >= 1 && l <= lmax - 1;
            if (!(interior || (isi ? a.m.per_i : a.m.per_j))) return;
            const int le = interior ? l : 0;
            const int nj = a.m.nj;
            const long long fidx = isi ? (long long)le * nj + t : (long long)t * (nj + 1) + le;
            g = load_geo(isi ? a.m.fi : a.m.fj, isi ? a.m.fi64 : a.m.fj64, isi ? a.m.nfi : a.m.nfj, fidx, 0.0, 0.0);
        }
    }

    HC_HD static void faces(const StageArgs<RF>& a, Smem& s, int i0, int j0, int tid, unsigned int& flag,
                            const Geo& pre)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        const QT* qs = &s.q[0][0][0];
        const RF* Qs = &s.Q[0][0][0][0];
        for (int k = tid; k < NF; k += NT) {
            const bool isi = face_kind_i(k);
            int li, lj;
            face_pos(k, isi, li, lj);
            const int l = isi ? i0 + li : j0 + lj;        // node line of the face
            const int t = isi ? j0 + lj : i0 + li;        // transverse cell index
            const int lmax = isi ? ni : nj;
            if (l > lmax || t >= (isi ? nj : ni)) continue;
            const bool interior = l >= 1 && l <= lmax - 1;
            const RF* base = isi ? a.m.fi : a.m.fj;
            const double* base64 = isi ? a.m.fi64 : a.m.fj64;
            const long long nf = isi ? a.m.nfi : a.m.nfj;
            const int pr = (li + 2) * QW + lj + 2;        // right cell in the q tile
            const int pl = pr - (isi ? QW : 1);            // left cell
            QT ql[NV], qr[NV];
            if (interior || (isi ? a.m.per_i : a.m.per_j)) {
                const int le = interior ? l : 0;
                const long long fidx = isi ? (long long)le * nj + t : (long long)t * (nj + 1) + le;
                const int mr = (li + 1) * MW + lj + 1;     // right cell in the Q ring
                const int ml = mr - (isi ? MW : 1);
                RF Ql[NV][2], Qr[NV][2];
                for (int v = 0; v < NV; ++v) {
                    ql[v] = qs[v * QH * QW + pl];
                    qr[v] = qs[v * QH * QW + pr];
                    Ql[v][0] = Qs[(2 * v) * MH * MW + ml];
                    Ql[v][1] = Qs[(2 * v + 1) * MH * MW + ml];
                    Qr[v][0] = Qs[(2 * v) * MH * MW + mr];
                    Qr[v][1] = Qs[(2 * v + 1) * MH * MW + mr];
                }
                double cx0 = 0.0, cx1 = 0.0;
                if constexpr (CSTAGE) centre_diff(a, s, pl, pr, !interior, isi, cx0, cx1);
                if constexpr (UNI != 0) {
                    store(s, k, muscl(a, uni_geo(a, isi), ql, qr, Ql, Qr));
                } else if (GPF != 0 && k == tid) {
                    Geo g = pre;
                    if constexpr (CSTAGE) {
                        g.dx0 = RF(cx0);
                        g.dx1 = RF(cx1);
                    }
                    store(s, k, muscl(a, g, ql, qr, Ql, Qr));
                } else {
                    store(s, k, muscl(a, load_geo(base, base64, nf, fidx, cx0, cx1), ql, qr, Ql, Qr));
                }
            } else {
                // the interior cell is right of node line 0 and left of the last node line
                const bool first = l == 0;
                const int bid = (isi ? 0 : 2) + (first ? 0 : 1);
                const int nc = first ? 0 : lmax - 1;
                const int pc = first ? pr : pl;
                for (int v = 0; v < NV; ++v) ql[v] = qs[v * QH * QW + pc];
                const long long fidx = isi ? (long long)l * nj + t : (long long)t * (nj + 1) + l;
                const long long cint = isi ? (long long)nc * nj + t : (long long)t * nj + nc;
                const long long row = isi ? (first ? 0 : nj) + t : 2LL * nj + (first ? 0 : ni) + t;
                store(s, k, boundary_face(a, bid, base, base64, nf, fidx, ql, cint, row, flag));
            }
        }
    }

    // ---- phase 4: owner-computes residual accumulation and the stage update ----
    // red: the cell's lambda/vol for the global-dt reduction of stage 0 (0 otherwise)
    // Cell prefetch: the cell phase's inputs of the lane's cell (step state, the accumulator read
    // first, volume, stage-0 spectral radius) are loaded into registers when the face phase opens,
    // so the cell phase does not start with a round trip to memory.  The 20 registers it holds
    // through the face phase decide the occupancy of the FP64 7x16 kernels: without it the
    // central, Rusanov, Roe-family and SLAU kernels, and AUSM+up in the conserved basis, fit 96
    // VGPRs (five waves per SIMD, which the 31 KB of LDS per work-group also allows), while
    // AUSM+up in the primitive basis needs more than 96 either way (four waves), so only that
    // kernel keeps the early prefetch (the 4096^2 vortex runs 3 % slower with the late one).
    static constexpr bool CPF = GPF != 0 && sizeof(RF) == sizeof(double) && LAW == LAW_EULER &&
                                FLUX == FLUX_AUSMPUP && BASIS == BASIS_PRIMITIVE;
    // The other FP64 prefetching configurations, and every configuration of the mixed objects but
    // AUSM+up's, load the same inputs at the end of the face phase, once the lane's face is stored:
    // the face phase's registers are free by then, and the loads are in flight while the work-group
    // waits at the barrier before the cell phase (the mixed AUSM+up kernels, FP32 faces throughout,
    // run 3 % slower with it).  The mixed objects are built with FMA contraction, so the choice is
    // made per flux for all tile configurations alike: a cell phase of another shape in some of them
    // could contract the update differently, and cases would then depend on the tile shape in the
    // last bits.
    static constexpr bool LPF = (GPF != 0 && !CPF) || (sizeof(RF) == sizeof(float) && FLUX != FLUX_AUSMPUP);
    static constexpr bool PF = CPF || LPF;
    struct CellIn {
        double q0[NV], acc[NV], vol, lam0;
    };

    // the accumulator value the cell phase reads first: the stage's base slot, else the old value
    // of its first update slot (-1: none)
    HC_HD static int first_acc_slot(const RkPlan& p)
    {
        if (p.base >= 0) return p.base;
        if (p.nupd > 0 && !p.upd_first[0]) return p.upd_slot[0];
        return -1;
    }

    HC_HD static void prefetch_cell(const StageArgs<RF>& a, int i0, int j0, int tid, CellIn& ci)
    {
        if constexpr (PF) {
            const int i = i0 + tid / TJ, j = j0 + tid % TJ;
            if (tid >= NC || i >= a.m.ni || j >= a.m.nj || a.mode == MODE_RESIDUAL) return;
            const long long c = (long long)i * a.m.nj + j;
            for (int v = 0; v < NV; ++v) ci.q0[v] = a.q0[v * a.N + c];
            const int fs = first_acc_slot(a.plan);
            if (fs >= 0)
                for (int v = 0; v < NV; ++v) ci.acc[v] = a.acc[((long long)fs * NV + v) * a.N + c];
            ci.vol = a.m.cvol[UNI ? 0 : c];
            if (a.dt_mode == 1 && a.stage != 0) ci.lam0 = a.lam0[c];
        }
    }

    HC_HD static void cells(const StageArgs<RF>& a, const Smem& s, int i0, int j0, int tid, unsigned int& flag,
                            double& red, const CellIn& ci)
    {
        const int ni = a.m.ni, nj = a.m.nj;
        const int li = tid / TJ, lj = tid % TJ;
        const int i = i0 + li, j = j0 + lj;
        red = 0.0;
        if (tid >= NC || i >= ni || j >= nj) return;
        double R[NV];
        double lam = 0.0;
        for (int v = 0; v < NV; ++v) R[v] = 0.0;
        auto plus = [&](int k) {
            for (int v = 0; v < NV; ++v) R[v] = R[v] + double(s.F[v][k]);
            lam = lam + double(s.F[NV][k]);
        };
        auto minus = [&](int k) {
            for (int v = 0; v < NV; ++v) R[v] = R[v] - double(s.F[v][k]);
            lam = lam + double(s.F[NV][k]);
        };
        const int kil = li * TJ + lj, kir = kil + TJ;                   // i-faces left / right
        const int kjl = NFI + li * (TJ + 1) + lj, kjr = kjl + 1;       // j-faces left / right
        // interior faces (residualCalc.h:199-253)
        if (i > 0) plus(kil);
        if (i < ni - 1) minus(kir);
        if (j > 0) plus(kjl);
        if (j < nj - 1) minus(kjr);
        // boundaries in id order (residualCalc.h:301-305)
        if (i == 0) {
            plus(kil);   // periodic wrap (+F on cell 0), ghost face (+F), wall (+F)
        }
        if (i == ni - 1) {
            if (a.bc[1] == BC_WALL) plus(kir);
            else minus(kir);   // periodic wrap (-F on cell ni-1), ghost face (-F)
        }
        if (j == 0) {
            plus(kjl);
        }
        if (j == nj - 1) {
            if (a.bc[3] == BC_WALL) plus(kjr);
            else minus(kjr);
        }
        const long long c = (long long)i * nj + j;
        if (a.mode == MODE_RESIDUAL || a.keep_res) {
            for (int v = 0; v < NV; ++v) a.res_out[c * (NV + 1) + v] = R[v];
            a.res_out[c * (NV + 1) + NV] = lam;
            if (a.mode == MODE_RESIDUAL) return;
        }
        // Runge-Kutta accumulation (rungeKutta.h:176-197)
        const RkPlan& p = a.plan;
        double Rt[NV];
        for (int v = 0; v < NV; ++v) {
            double base = 0.0;
            if (p.base >= 0) base = PF ? ci.acc[v] : a.acc[((long long)p.base * NV + v) * a.N + c];
            Rt[v] = base + R[v] * p.a_ss;
        }
        // unrolled over the plan's fixed capacity: constant indices only (the arguments may live in registers)
#pragma unroll
        for (int u = 0; u < 5; ++u) {
            if (u >= p.nupd) break;
            for (int v = 0; v < NV; ++v) {
                double* A = &a.acc[((long long)p.upd_slot[u] * NV + v) * a.N + c];
                double old = 0.0;
                if (!p.upd_first[u]) old = (PF && u == 0 && p.base < 0) ? ci.acc[v] : *A;
                *A = old + R[v] * p.upd_coef[u];
            }
        }
        const double vol = PF ? ci.vol : a.m.cvol[UNI ? 0 : c];
        if (a.mode == MODE_GLOBAL_STAGE0) {
            for (int v = 0; v < NV; ++v) a.tmp[v * a.N + c] = Rt[v];
            red = lam / vol;   // spectralRadius.h:33-35
            return;
        }
        double fac;
        if (a.dt_mode == 0) {
            const double dt = a.dtslot[0];
            fac = (p.beta * dt) / vol;
        } else if (a.dt_mode == 2) {
            fac = (p.beta * a.cfl) / vol;   // fixed step: cfl holds dt
        } else {
            double l0;
            if (a.stage == 0) {
                l0 = lam;
                a.lam0[c] = lam;
            } else {
                l0 = PF ? ci.lam0 : a.lam0[c];
            }
            fac = (p.beta * a.cfl) / l0;
        }
        double q0[NV];
        for (int v = 0; v < NV; ++v) q0[v] = PF ? ci.q0[v] : a.q0[v * a.N + c];
        update_cell(a, c, q0, Rt, fac, flag);
    }

    // q_out = S( C(q0) + R_tot*fac )   (spatial/eulerForwardUpdate.h:76-85)
    HC_HD static void update_cell(const StageArgs<RF>& a, long long c, const double* q0, const double* Rt, double fac,
                                  unsigned int& flag)
    {
        double qc[NV], qn[NV];
        to_conserved<double, LAW, BASIS>(a.sp, q0, qc);
        for (int v = 0; v < NV; ++v) qc[v] = qc[v] + Rt[v] * fac;
        from_conserved<double, LAW, BASIS>(a.sp, qc, qn);
        bool fin = true;
        for (int v = 0; v < NV; ++v) {
            a.q_out[v * a.N + c] = qn[v];
            fin = fin && hc_isfinite(qn[v]);
        }
        if (!fin) flag |= ST_NONFINITE;
        if constexpr (LAW == LAW_EULER && BASIS == BASIS_PRIMITIVE) {
            if (!(qn[2] > 0.0) || !(qn[3] > 0.0)) flag |= ST_NONPHYSICAL;
        } else if constexpr (LAW == LAW_EULER) {
            if (!(qn[2] > 0.0)) flag |= ST_NONPHYSICAL;
        }
    }
};

// deferred update of global-dt stage 0 (dt = cfl / max(lambda/vol), rungeKutta.h:133)
template <int LAW, int BASIS> struct Stage0Update {
    static constexpr int NV = LawTraits<LAW>::nv;
    // mx: the case's max(lambda/vol) over its slots
    template <class RF>
    HC_HD static void run(const StageArgs<RF>& a, long long c, unsigned int& flag, double mx)
    {
        const double dt = a.cfl / mx;
        if (c == 0) {
            a.dtslot[0] = dt;
            a.dtslot[1] = a.dtslot[1] + dt;
        }
        if (c >= a.N) return;
        const double stepfac = a.plan.beta * dt;
        const double fac = stepfac / a.m.cvol[c];
        double Rt[NV], q0[NV];
        for (int v = 0; v < NV; ++v) {
            Rt[v] = a.tmp[v * a.N + c];
            q0[v] = a.q0[v * a.N + c];
        }
        StageTile<LAW, BASIS, FLUX_CENTRAL, RF, 1, 64>::update_cell(a, c, q0, Rt, fac, flag);
    }
};

// ---------------------------------------------------------------------------------------
// Batched launch: one launch per stage covers the tiles of every case of a context.  The
// per-case arrays are packed one case after the other (state [nv][N_c] blocks, face slots
// [NSLOT][nf_c] blocks; cases whose geometry is bitwise equal share one geometry block); a case
// record holds the case's mesh view with its pointers already offset, its species, boundary types
// and offsets into the packed per-cell arrays.
template <class RF> struct CaseRec {
    MeshView<RF> m;
    Species<double> sp;
    Species<RF> spf;
    int bc[4];
    long long N;          // cells of the case
    long long cell_off;   // first cell of the case in the packed per-cell arrays
    long long bref_off;   // first boundary-reference row of the case
    int tiles_j;          // tiles along j in the record's tile shape
};

// One entry of a context's tile table: the case and the tile within the case (row-major over the
// case's TI x TJ tiles).  Cases that share their geometry are interleaved tile by tile in the
// table, so the work-groups an XCD runs together read the same face slots (one HBM fetch, the
// rest L2 hits); the table also holds every case's tiles contiguously for single-case launches.
struct TileRef {
    int ci;
    int lt;
};

template <class RF> struct LaunchArgs {
    const CaseRec<RF>* cases;
    const TileRef* tiles;   // the context's tile table
    int tile_base;          // first entry of the launch in the table
    int ntiles;             // tiles of the launch
    int nslots;
    int vscale, pscale, limiter;
    int mode, dt_mode, stage, reset_max;
    int keep_res;         // fused stages: also store the stage's residual in res_out (stage 0 of a step)
    double cfl;
    RkPlan plan;
    const double* q_in;   // packed [nv][N_c] blocks
    const double* q0;
    double* q_out;
    double* acc;          // packed [nslots][nv][N_c] blocks
    double* tmp;
    double* lam0;
    const double* bref;   // packed [rows_c][nv] blocks
    double* res_out;      // packed [N_c][nv+1] blocks
    unsigned long long* maxslot;   // MAXSPREAD * MAXSTRIDE per case
    double* dtslot;                // two per case
    unsigned int* status;
};

// the single-case view of case ci of a batched launch
template <class RF, int NV> HC_HD StageArgs<RF> case_args(const LaunchArgs<RF>& L, const CaseRec<RF>& c, int ci)
{
    StageArgs<RF> a;
    a.m = c.m;
    a.sp = c.sp;
    a.spf = c.spf;
    a.vscale = L.vscale;
    a.pscale = L.pscale;
    a.limiter = L.limiter;
    for (int b = 0; b < 4; ++b) a.bc[b] = c.bc[b];
    a.mode = L.mode;
    a.dt_mode = L.dt_mode;
    a.stage = L.stage;
    a.reset_max = L.reset_max;
    a.keep_res = L.keep_res;
    a.N = c.N;
    a.cfl = L.cfl;
    a.q_in = L.q_in + NV * c.cell_off;
    a.q0 = L.q0 + NV * c.cell_off;
    a.q_out = L.q_out + NV * c.cell_off;
    a.acc = L.nslots > 0 ? L.acc + (long long)L.nslots * NV * c.cell_off : nullptr;
    a.tmp = L.tmp + NV * c.cell_off;
    a.lam0 = L.lam0 + c.cell_off;
    a.bref = L.bref + NV * c.bref_off;
    a.res_out = L.res_out + (NV + 1) * c.cell_off;
    a.maxslot = L.maxslot + (long long)ci * MAXSPREAD * MAXSTRIDE;
    a.dtslot = L.dtslot + 2 * ci;
    a.status = L.status;
    a.plan = L.plan;
    return a;
}

// chunk of cells of the deferred global-dt update: 256 consecutive cells of one case
struct UpdChunk {
    int ci;
    int start;
};

#if defined(__HIPCC__)

// wave-level reductions for gfx950's 64-lane wavefronts
__device__ inline double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ inline unsigned int wave_or(unsigned int v)
{
    for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
    return v;
}
__device__ inline unsigned long long wave_max_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}
static_assert(MAXSPREAD == 64, "one slot per lane of a wavefront");

// optional occupancy hint (waves per SIMD) for the register allocator, set per build
#if defined(HC_STAGE_WAVES_PER_EU) && HC_STAGE_WAVES_PER_EU > 0
#define HC_STAGE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(HC_STAGE_WAVES_PER_EU)))
#else
#define HC_STAGE_OCCUPANCY
#endif

// The prefetching configurations hold the prefetched values across the metric phase: at least 4
// waves per SIMD keeps them within 128 VGPRs (the AUSM+up FP64 kernel would otherwise take 129
// and drop to 3 waves)
template <int LAW, int BASIS, int FLUX, class RF, int TI, int TJ, int NTHR, int CST, int GPF, int UNI = 0>
__global__ void __launch_bounds__(NTHR) HC_STAGE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(GPF != 0 ? 4 : 1)))
k_stage(LaunchArgs<RF> L)
{
    using T = StageTile<LAW, BASIS, FLUX, RF, TI, TJ, NTHR, CST, GPF, UNI>;
    __shared__ typename T::Smem s;
    const int tid = threadIdx.x;
    // XCD-aware tile order: consecutive work-group ids are dealt round-robin to the
    // 8 XCDs, so give each XCD a contiguous band of tiles (shared halos stay in one L2)
    const TileRef tr = L.tiles[L.tile_base + xcd_tile(blockIdx.x, L.ntiles)];
    const int ci = tr.ci, lt = tr.lt;
    const CaseRec<RF>& c = L.cases[ci];
    const int i0 = (lt / c.tiles_j) * TI, j0 = (lt % c.tiles_j) * TJ;
    // The case's parameters (18 array pointers, species, Runge-Kutta plan) held from the first
    // phase to the last exceed the scalar register file, and the kernels then spill them into
    // vector-register lanes (about 110 lane moves per wavefront).  Instead, a group of phases reads
    // them again from the case record and the kernel arguments (scalar-cache hits) through pointers
    // passed through an empty asm, which hides that they are the ones read before, so the values
    // of the previous group die at its barrier: the general-geometry kernels read them for the
    // face phase and again for the cell phase, the uniform-geometry kernels once more for the cell
    // phase only (a read for their face phase costs them vector registers and time).
    auto args = [&](bool reread = UNI == 0) -> StageArgs<RF> {
        if (!reread) {
            return case_args<RF, T::NV>(L, c, ci);
        } else {
            using ConstRec = const __attribute__((address_space(4))) CaseRec<RF>;
            using ConstArgs = const __attribute__((address_space(4))) LaunchArgs<RF>;
            ConstRec* cr = (ConstRec*)&L.cases[__builtin_amdgcn_readfirstlane(ci)];
            ConstArgs* la = (ConstArgs*)__builtin_amdgcn_kernarg_segment_ptr();   // L: the only argument, at offset 0
            asm volatile("" : "+s"(cr), "+s"(la) : : "memory");
            return case_args<RF, T::NV>(*(const LaunchArgs<RF>*)la, *(const CaseRec<RF>*)cr, ci);
        }
    };
    unsigned int flag = 0;
    typename T::Geo g0;
    typename T::CellIn cin;
    {
        const StageArgs<RF> a = args();
        T::load(a, s, i0, j0, tid);
        T::prefetch(a, i0, j0, tid, g0);
        __syncthreads();
        T::qmetric(a, s, i0, j0, tid);
    }
    __syncthreads();
    {
        const StageArgs<RF> a = args();
        if constexpr (T::CPF) T::prefetch_cell(a, i0, j0, tid, cin);
        T::faces(a, s, i0, j0, tid, flag, g0);
        if constexpr (T::LPF) T::prefetch_cell(a, i0, j0, tid, cin);
    }
    __syncthreads();
    double red;
    const StageArgs<RF> a = args(true);
    T::cells(a, s, i0, j0, tid, flag, red, cin);
    // reductions: the status bits; for global-dt stage 0 also max(lambda/vol), over the
    // work-group and into the tile's slot (mode is uniform over the grid: the barrier is too)
    const double m = wave_max(red);
    const unsigned int f = wave_or(flag);
    if (a.mode == MODE_GLOBAL_STAGE0) {
        static_assert(T::NT % 64 == 0, "whole wavefronts");
        __shared__ unsigned long long wm[T::NT / 64];
        __shared__ unsigned int wf[T::NT / 64];
        if ((tid & 63) == 0) {
            wm[tid >> 6] = dbl_bits(m);
            wf[tid >> 6] = f;
        }
        __syncthreads();
        if (tid == 0) {
            unsigned long long mm = wm[0];
            unsigned int ff = wf[0];
            for (int w = 1; w < T::NT / 64; ++w) {
                mm = wm[w] > mm ? wm[w] : mm;
                ff |= wf[w];
            }
            atomicMax(a.maxslot + (lt % MAXSPREAD) * MAXSTRIDE, mm);
            if (ff) atomicOr(a.status, ff);
        }
    } else if ((tid & 63) == 0 && f) {
        atomicOr(a.status, f);
    }
    // the last stage re-arms the case's max slots for the next step (no stage reads them then)
    if (a.reset_max && lt == 0 && tid < MAXSPREAD) a.maxslot[tid * MAXSTRIDE] = dbl_bits(DBL_MIN);
    // fixed step: the case's time advances once per step (one thread of the case's first tile)
    if (a.dt_mode == 2 && a.mode == MODE_FUSED && a.stage == 0 && lt == 0 && tid == 0) {
        a.dtslot[0] = a.cfl;
        a.dtslot[1] = a.dtslot[1] + a.cfl;
    }
}

template <int LAW, int BASIS, class RF>
__global__ void __launch_bounds__(256) k_stage0_update(LaunchArgs<RF> L, const UpdChunk* chunks)
{
    const UpdChunk ch = chunks[blockIdx.x];
    const StageArgs<RF> a = case_args<RF, LawTraits<LAW>::nv>(L, L.cases[ch.ci], ch.ci);
    const long long c = (long long)ch.start + threadIdx.x;
    // every lane is active here: one slot per lane, then the wavefront's max
    const double mx = bits_dbl(wave_max_u64(a.maxslot[(threadIdx.x & 63) * MAXSTRIDE]));
    unsigned int flag = 0;
    Stage0Update<LAW, BASIS>::run(a, c, flag, mx);
    const unsigned int f = wave_or(flag);
    if ((threadIdx.x & 63) == 0 && f) atomicOr(a.status, f);
}

#endif

}   // namespace hc
