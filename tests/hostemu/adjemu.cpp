// adjemu.cpp -- TEST-ONLY host execution of the tangent and adjoint residual kernels
// (hc_adjoint.hpp), face by face and cell by cell, with the product's geometry.  Tests use it
// to check the derivatives against finite differences of the oracle and the tangent/adjoint
// pair against each other without a GPU.  Never part of the product library.
#include "hc_adjoint.hpp"
#include "hc_geom.hpp"
#include "hcamd.h"

#include <vector>

using namespace hc;

namespace {

struct AdjEmu {
    GeomOutput g;
    std::vector<double> q, qd, bref, brefd, Q, Qd, F, Fd, res, resd, ybar, minf_face, brefbar, qbar, Qbar;
    AdjArgs a{};
};

int setup(AdjEmu& e, const hc_case_desc* cs, int ni, int nj, const double* nodes, const double* q, const double* bref)
{
    GeomInput gin{ni, nj, nodes, {cs->bc[0], cs->bc[1], cs->bc[2], cs->bc[3]}, {cs->period_i[0], cs->period_i[1]},
                  {cs->period_j[0], cs->period_j[1]}};
    if (build_geometry(gin, e.g) != 0) return -1;
    const int nv = cs->law == HC_LAW_EULER ? 4 : 3;
    const long long N = (long long)ni * nj;
    AdjArgs& a = e.a;
    a.m.ni = ni;
    a.m.nj = nj;
    a.m.per_i = cs->bc[0] == HC_BC_PERIODIC;
    a.m.per_j = cs->bc[2] == HC_BC_PERIODIC;
    a.m.nfi = (long long)(ni + 1) * nj;
    a.m.nfj = (long long)ni * (nj + 1);
    a.m.cvol = e.g.cvol.data();
    a.m.ccx = e.g.ccx.data();
    a.m.ccy = e.g.ccy.data();
    a.m.fi = e.g.fi.data();
    a.m.fj = e.g.fj.data();
    a.m.fi64 = e.g.fi.data();
    a.m.fj64 = e.g.fj.data();
    const hc_species& s = cs->sp;
    a.sp = Species<double>{s.gamma, s.minf, s.lref, s.nu, s.pr, s.dt, s.R, s.gamma1};
    a.vscale = cs->vscale;
    a.pscale = cs->pscale;
    a.limiter = cs->limiter;
    for (int b = 0; b < 4; ++b) a.bc[b] = cs->bc[b];
    a.N = N;
    a.nfaces = a.m.nfi + a.m.nfj;
    e.q.resize((size_t)nv * N);
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < nv; ++v) e.q[(size_t)v * N + c] = q[c * nv + v];
    e.bref.assign(bref, bref + (size_t)(2 * nj + 2 * ni) * nv);
    e.Q.assign((size_t)2 * nv * N, 0.0);
    a.q = e.q.data();
    a.bref = e.bref.data();
    a.Q = e.Q.data();
    return nv;
}

template <int LAW, int BASIS, int FLUX> void run_tangent(AdjArgs& a)
{
    using K = AdjKernels<LAW, BASIS, FLUX>;
    for (long long c = 0; c < a.N; ++c) K::tan_qmetric(a, c);
    for (long long f = 0; f < a.nfaces; ++f) K::tan_face(a, f);
    for (long long c = 0; c < a.N; ++c) K::tan_cell(a, c);
}

// the GPU's launches in order: the tiles (each phase for every thread of the work-group before the
// next), the metric transpose
template <int LAW, int BASIS, int FLUX> void run_adjoint(AdjArgs& a)
{
    using K = AdjKernels<LAW, BASIS, FLUX>;
    using T = AdjTile<LAW, BASIS, FLUX, ADJ_TI, ADJ_TJ, ADJ_NT>;
    std::vector<typename T::Smem> smem(1);
    const int tiles_i = (a.m.ni + ADJ_TI - 1) / ADJ_TI, tiles_j = (a.m.nj + ADJ_TJ - 1) / ADJ_TJ;
    for (int t = 0; t < tiles_i * tiles_j; ++t) {
        const int i0 = (t / tiles_j) * ADJ_TI, j0 = (t % tiles_j) * ADJ_TJ;
        for (int tid = 0; tid < ADJ_NT; ++tid) T::load(a, smem[0], i0, j0, tid);
        for (int tid = 0; tid < ADJ_NT; ++tid) T::qmetric(a, smem[0], i0, j0, tid);
        for (int tid = 0; tid < ADJ_NT; ++tid) T::faces(a, smem[0], i0, j0, tid);
        for (int tid = 0; tid < ADJ_NT; ++tid) T::cells(a, smem[0], i0, j0, tid);
    }
    for (long long c = 0; c < a.N; ++c) K::adj_qedges(a, c);
}

#define ADJ_KEYS(X) X(1, 0, 0) X(1, 0, 1) X(1, 0, 2) X(1, 0, 3) X(1, 0, 4) X(1, 0, 5) X(1, 0, 6) X(1, 1, 4)

bool dispatch(const hc_case_desc* cs, AdjArgs& a, bool tangent)
{
    const int key = cs->law * 100 + cs->basis * 10 + cs->flux;
#define ADJ_CASE(L, B, F)                                  \
    case L * 100 + B * 10 + F:                             \
        if (tangent) run_tangent<L, B, F>(a);              \
        else run_adjoint<L, B, F>(a);                      \
        return true;
    switch (key) {
        ADJ_KEYS(ADJ_CASE)
    default: return false;
    }
#undef ADJ_CASE
}

}   // namespace

extern "C" {

// R(q) and Rdot for the tangent (qd, brefd, minfd); q, qd AoS [N][nv]; res [N][nv+1]; resd [N][nv]
int emu_tangent_residual(const hc_case_desc* cs, int ni, int nj, const double* nodes, const double* q, const double* qd,
                         const double* bref, const double* brefd, double minfd, double* res, double* resd)
{
    AdjEmu e;
    const int nv = setup(e, cs, ni, nj, nodes, q, bref);
    if (nv < 0) return -1;
    const long long N = e.a.N;
    e.qd.resize((size_t)nv * N);
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < nv; ++v) e.qd[(size_t)v * N + c] = qd[c * nv + v];
    if (brefd) e.brefd.assign(brefd, brefd + e.bref.size());
    e.Qd.assign(e.Q.size(), 0.0);
    e.F.assign((size_t)(nv + 1) * e.a.nfaces, 0.0);
    e.Fd.assign((size_t)nv * e.a.nfaces, 0.0);
    e.a.qd = e.qd.data();
    e.a.brefd = brefd ? e.brefd.data() : nullptr;
    e.a.minfd = minfd;
    e.a.Qd = e.Qd.data();
    e.a.F = e.F.data();
    e.a.Fd = e.Fd.data();
    e.a.res = res;
    e.a.resd = resd;
    return dispatch(cs, e.a, true) ? 0 : -2;
}

// qbar = (dR/dq)^T ybar, brefbar, minfbar; ybar, qbar AoS [N][nv]; brefbar [rows][nv]
int emu_adjoint_residual(const hc_case_desc* cs, int ni, int nj, const double* nodes, const double* q,
                         const double* bref, const double* ybar, double* qbar, double* brefbar, double* minfbar)
{
    AdjEmu e;
    const int nv = setup(e, cs, ni, nj, nodes, q, bref);
    if (nv < 0) return -1;
    const long long N = e.a.N;
    e.ybar.resize((size_t)nv * N);
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < nv; ++v) e.ybar[(size_t)v * N + c] = ybar[c * nv + v];
    e.minf_face.assign((size_t)e.a.nfaces, 0.0);
    e.brefbar.assign(e.bref.size(), 0.0);
    e.qbar.assign((size_t)nv * N, 0.0);
    e.Qbar.assign((size_t)2 * nv * N, 0.0);
    e.a.ybar = e.ybar.data();
    e.a.minf_face = e.minf_face.data();
    e.a.brefbar = e.brefbar.data();
    e.a.qbar = e.qbar.data();
    e.a.Qbar = e.Qbar.data();
    if (!dispatch(cs, e.a, false)) return -2;
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < nv; ++v) qbar[c * nv + v] = e.qbar[(size_t)v * N + c];
    for (size_t k = 0; k < e.brefbar.size(); ++k) brefbar[k] = e.brefbar[k];
    double m = 0.0;
    for (long long f = 0; f < e.a.nfaces; ++f) m = m + e.minf_face[f];
    *minfbar = m;
    return 0;
}

}   // extern "C"
