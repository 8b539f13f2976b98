// hostemu.cpp -- TEST-ONLY host emulation of the stage kernels.
//
// Executes the HIP stage kernel's phase functions (hc_stage.hpp) work-group by
// work-group and thread by thread on the CPU, with the same geometry (hc_geom.cpp),
// buffer rotation and Runge-Kutta plans as the C-ABI.  Tests use it to check the
// kernels' arithmetic and LDS indexing against the oracle without a GPU.  It is
// never part of the product library.
#include "hc_geom.hpp"
#include "hc_stage.hpp"
#include "hc_tiles.hpp"
#include "hcamd.h"

#include <cfloat>
#include <memory>
#include <vector>

using namespace hc;

namespace {

struct Emu {
    int ni, nj, nv, precision;
    hc_case_desc cs;
    GeomOutput g;
    std::vector<float> fi32, fj32;
    std::vector<double> q[3];
    int cur = 0;
    std::vector<double> acc, tmp, lam0, bref, res;
    unsigned long long maxslot;
    double dtslot[2] = {0, 0};
    unsigned int status = 0;
    std::vector<RkPlan> plans;
    int nslots = 0;
};

template <class RF> StageArgs<RF> args_of(Emu& e)
{
    StageArgs<RF> a{};
    a.m.ni = e.ni;
    a.m.nj = e.nj;
    a.m.per_i = e.cs.bc[0] == BC_PERIODIC;
    a.m.per_j = e.cs.bc[2] == BC_PERIODIC;
    a.m.nfi = (long long)(e.ni + 1) * e.nj;
    a.m.nfj = (long long)e.ni * (e.nj + 1);
    a.m.cvol = e.g.cvol.data();
    a.m.ccx = e.g.ccx.data();
    a.m.ccy = e.g.ccy.data();
    if constexpr (sizeof(RF) == 8) {
        a.m.fi = e.g.fi.data();
        a.m.fj = e.g.fj.data();
    } else {
        a.m.fi = e.fi32.data();
        a.m.fj = e.fj32.data();
    }
    a.m.fi64 = e.g.fi.data();
    a.m.fj64 = e.g.fj.data();
    for (int d = 0; d < 2; ++d) {
        a.m.period_i[d] = e.cs.period_i[d];
        a.m.period_j[d] = e.cs.period_j[d];
    }
    const hc_species& s = e.cs.sp;
    a.sp = Species<double>{s.gamma, s.minf, s.lref, s.nu, s.pr, s.dt, s.R, s.gamma1};
    a.spf = Species<RF>{RF(s.gamma), RF(s.minf), RF(s.lref), RF(s.nu), RF(s.pr), RF(s.dt), RF(s.R), RF(s.gamma1)};
    a.vscale = e.cs.vscale;
    a.pscale = e.cs.pscale;
    a.limiter = e.cs.limiter;
    for (int b = 0; b < 4; ++b) a.bc[b] = e.cs.bc[b];
    a.dt_mode = e.cs.dt_mode;
    a.N = (long long)e.ni * e.nj;
    a.cfl = e.cs.cfl;
    a.acc = e.acc.data();
    a.tmp = e.tmp.data();
    a.lam0 = e.lam0.data();
    a.bref = e.bref.data();
    a.res_out = e.res.data();
    a.maxslot = &e.maxslot;
    a.dtslot = e.dtslot;
    a.status = &e.status;
    return a;
}

// uniform-geometry kernels (UNI) when the case allows them, selected with emu_set_uniform_geometry
int g_uniform = 0;

template <int LAW, int BASIS, int FLUX, class RF, int TI, int TJ, int NT, int CST, int GPF, int UNI>
void run_stage_t(const StageArgs<RF>& a)
{
    using T = StageTile<LAW, BASIS, FLUX, RF, TI, TJ, NT, CST, GPF, UNI>;
    const int tiles_i = (a.m.ni + TI - 1) / TI, tiles_j = (a.m.nj + TJ - 1) / TJ;
    const int ntiles = tiles_i * tiles_j;
    auto sp = std::make_unique<typename T::Smem>();
    typename T::Smem& s = *sp;
    std::vector<unsigned int> flag(T::NT);
    for (int b = 0; b < ntiles; ++b) {
        const int t = xcd_tile(b, ntiles);
        const int i0 = (t / tiles_j) * TI, j0 = (t % tiles_j) * TJ;
        for (auto& f : flag) f = 0;
        std::vector<typename T::Geo> g0(T::NT);
        for (int tid = 0; tid < T::NT; ++tid) {
            T::load(a, s, i0, j0, tid);
            T::prefetch(a, i0, j0, tid, g0[tid]);
        }
        for (int tid = 0; tid < T::NT; ++tid) T::qmetric(a, s, i0, j0, tid);
        for (int tid = 0; tid < T::NT; ++tid) T::faces(a, s, i0, j0, tid, flag[tid], g0[tid]);
        std::vector<double> red(T::NT);
        std::vector<typename T::CellIn> ci(T::NT);
        for (int tid = 0; tid < T::NT; ++tid) T::prefetch_cell(a, i0, j0, tid, ci[tid]);
        for (int tid = 0; tid < T::NT; ++tid) T::cells(a, s, i0, j0, tid, flag[tid], red[tid], ci[tid]);
        double m = 0.0;
        unsigned int f = 0;
        for (int tid = 0; tid < T::NT; ++tid) {
            m = fmax(m, red[tid]);
            f |= flag[tid];
        }
        if (a.mode == MODE_GLOBAL_STAGE0 && dbl_bits(m) > *a.maxslot) *a.maxslot = dbl_bits(m);
        *a.status |= f;
    }
    if (a.reset_max) *a.maxslot = dbl_bits(DBL_MIN);
}

bool g_case_uniform = false;

template <int LAW, int BASIS, int FLUX, class RF, int TI, int TJ, int NT, int CST, int GPF>
void run_stage(const StageArgs<RF>& a)
{
    if (g_uniform && g_case_uniform) run_stage_t<LAW, BASIS, FLUX, RF, TI, TJ, NT, 0, GPF, 1>(a);
    else run_stage_t<LAW, BASIS, FLUX, RF, TI, TJ, NT, CST, GPF, 0>(a);
}

template <int LAW, int BASIS, class RF> void run_update(const StageArgs<RF>& a)
{
    unsigned int flag = 0;
    // the device spreads the max over slots; its max over them is this one slot's value
    const double mx = bits_dbl(*a.maxslot);
    for (long long c = 0; c < a.N; ++c) Stage0Update<LAW, BASIS>::run(a, c, flag, mx);
    *a.status |= flag;
}

template <class RF, int TI, int TJ, int NT, int CST, int GPF> void dispatch_stage(const hc_case_desc& cs, const StageArgs<RF>& a)
{
    const int key = cs.law * 100 + cs.basis * 10 + cs.flux;
    switch (key) {
    case 0: run_stage<0, 0, 0, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 1: run_stage<0, 0, 1, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 100: run_stage<1, 0, 0, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 101: run_stage<1, 0, 1, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 102: run_stage<1, 0, 2, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 103: run_stage<1, 0, 3, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 104: run_stage<1, 0, 4, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 105: run_stage<1, 0, 5, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 106: run_stage<1, 0, 6, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 110: run_stage<1, 1, 0, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 111: run_stage<1, 1, 1, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 112: run_stage<1, 1, 2, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 113: run_stage<1, 1, 3, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 114: run_stage<1, 1, 4, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 115: run_stage<1, 1, 5, RF, TI, TJ, NT, CST, GPF>(a); break;
    case 116: run_stage<1, 1, 6, RF, TI, TJ, NT, CST, GPF>(a); break;
    default: break;
    }
}

template <class RF> void dispatch_update(const hc_case_desc& cs, const StageArgs<RF>& a)
{
    const int key = cs.law * 10 + cs.basis;
    if (key == 0) run_update<0, 0, RF>(a);
    else if (key == 10) run_update<1, 0, RF>(a);
    else run_update<1, 1, RF>(a);
}

// tile shape of the emulated launch (hc_tiles.hpp), selected with emu_set_tile_config
int g_tile_cfg = 0;

template <int K> void stage_shape(Emu& e, StageArgs<double>& a64, StageArgs<float>& a32)
{
    constexpr TileShape S64 = TILES64[K], S32 = TILES32[K];
    if (e.precision != HC_PRECISION_MIXED) dispatch_stage<double, S64.ti, S64.tj, S64.nt, S64.cst, S64.gpf>(e.cs, a64);
    else if (mixed_wide(e.cs.law, e.cs.flux))
        dispatch_stage<float, S64.ti, S64.tj, S64.nt, S64.cst, S64.gpf>(e.cs, a32);
    else dispatch_stage<float, S32.ti, S32.tj, S32.nt, S32.cst, S32.gpf>(e.cs, a32);
}

void stage(Emu& e, StageArgs<double>& a64, StageArgs<float>& a32)
{
    static_assert(NTILECFG == 3, "one stage_shape per tile configuration");
    if (g_tile_cfg == 2) stage_shape<2>(e, a64, a32);
    else if (g_tile_cfg == 1) stage_shape<1>(e, a64, a32);
    else stage_shape<0>(e, a64, a32);
}

void update(Emu& e, StageArgs<double>& a64, StageArgs<float>& a32)
{
    if (e.precision == HC_PRECISION_MIXED) dispatch_update<float>(e.cs, a32);
    else dispatch_update<double>(e.cs, a64);
}

int setup(Emu& e, const hc_case_desc* cs, int ni, int nj, const double* nodes, const double* q, const double* bref,
          int precision)
{
    e.ni = ni;
    e.nj = nj;
    e.cs = *cs;
    e.nv = cs->law == HC_LAW_EULER ? 4 : 3;
    e.precision = precision;
    GeomInput gin{ni, nj, nodes, {cs->bc[0], cs->bc[1], cs->bc[2], cs->bc[3]}, {cs->period_i[0], cs->period_i[1]},
                  {cs->period_j[0], cs->period_j[1]}};
    if (build_geometry(gin, e.g) != 0) return -1;
    g_case_uniform = uniform_geometry(gin, e.g);
    fp32_slots(e.g.fi, (long long)(ni + 1) * nj, e.fi32);
    fp32_slots(e.g.fj, (long long)ni * (nj + 1), e.fj32);
    const long long N = (long long)ni * nj;
    for (auto& b : e.q) b.assign((size_t)e.nv * N, 0.0);
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < e.nv; ++v) e.q[0][(size_t)v * N + c] = q[c * e.nv + v];
    e.nslots = make_rk_plans(cs->nstages, cs->alpha, cs->beta, e.plans);
    e.acc.assign((size_t)(e.nslots > 0 ? e.nslots : 1) * e.nv * N, 0.0);
    e.tmp.assign((size_t)e.nv * N, 0.0);
    e.lam0.assign(N, 0.0);
    e.res.assign((size_t)(e.nv + 1) * N, 0.0);
    e.bref.assign(bref, bref + (size_t)(2 * nj + 2 * ni) * e.nv);
    e.maxslot = dbl_bits(DBL_MIN);
    return 0;
}

}   // namespace

extern "C" {

// 1: run the uniform-geometry kernels on cases whose geometry allows them; returns whether the
// last set-up case allowed them
int emu_set_uniform_geometry(int enable)
{
    g_uniform = enable;
    return g_case_uniform ? 1 : 0;
}

int emu_set_tile_config(int cfg)
{
    if (cfg < 0 || cfg >= NTILECFG) return -1;
    g_tile_cfg = cfg;
    return 0;
}

int emu_residual(const hc_case_desc* cs, int ni, int nj, const double* nodes, const double* q, const double* bref,
                 int precision, double* res, unsigned int* status)
{
    Emu e;
    if (setup(e, cs, ni, nj, nodes, q, bref, precision)) return -1;
    StageArgs<double> a64 = args_of<double>(e);
    StageArgs<float> a32 = args_of<float>(e);
    a64.mode = a32.mode = MODE_RESIDUAL;
    a64.q_in = a32.q_in = e.q[0].data();
    a64.q0 = a32.q0 = e.q[0].data();
    a64.plan = a32.plan = e.plans[0];
    stage(e, a64, a32);
    const long long N = (long long)ni * nj;
    for (long long k = 0; k < (e.nv + 1) * N; ++k) res[k] = e.res[k];
    *status = e.status;
    return 0;
}

// the launch sequence of hc_capi.cpp enqueue_step, executed on the host
int emu_integrate(const hc_case_desc* cs, int ni, int nj, const double* nodes, double* q, const double* bref,
                  int precision, int nsteps, double* t_out, unsigned int* status)
{
    Emu e;
    if (setup(e, cs, ni, nj, nodes, q, bref, precision)) return -1;
    const int ns = cs->nstages;
    const bool global = cs->dt_mode == HC_DT_GLOBAL;
    for (int step = 0; step < nsteps; ++step) {
        double* q0 = e.q[e.cur].data();
        double* bufA = ns == 1 ? e.q[1 - e.cur].data() : e.q[(e.cur + 1) % 3].data();
        double* bufB = ns == 1 ? bufA : e.q[(e.cur + 2) % 3].data();
        StageArgs<double> a64 = args_of<double>(e);
        StageArgs<float> a32 = args_of<float>(e);
        const double* in = q0;
        for (int s = 0; s < ns; ++s) {
            double* out = ns == 1 ? bufA : (s == ns - 1 ? q0 : (s % 2 == 0 ? bufA : bufB));
            a64.q_in = a32.q_in = in;
            a64.q0 = a32.q0 = q0;
            a64.q_out = a32.q_out = out;
            a64.stage = a32.stage = s;
            a64.plan = a32.plan = e.plans[s];
            a64.reset_max = a32.reset_max = (global && ns >= 2 && s == ns - 1) ? 1 : 0;
            a64.mode = a32.mode = (global && s == 0) ? MODE_GLOBAL_STAGE0 : MODE_FUSED;
            stage(e, a64, a32);
            if (global && s == 0) update(e, a64, a32);
            in = out;
        }
        if (global && ns == 1) e.maxslot = dbl_bits(DBL_MIN);
        if (cs->dt_mode == HC_DT_FIXED) {
            e.dtslot[0] = cs->cfl;
            e.dtslot[1] = e.dtslot[1] + cs->cfl;
        }
        if (ns == 1) e.cur = 1 - e.cur;
    }
    const long long N = (long long)ni * nj;
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < e.nv; ++v) q[c * e.nv + v] = e.q[e.cur][(size_t)v * N + c];
    *t_out = e.dtslot[1];
    *status = e.status;
    return 0;
}

// geometry of the product (cells, LSQ metric, face slots) for inspection
int emu_geometry(const hc_case_desc* cs, int ni, int nj, const double* nodes, double* cells3, double* X3, double* fi,
                 double* fj)
{
    GeomOutput g;
    GeomInput gin{ni, nj, nodes, {cs->bc[0], cs->bc[1], cs->bc[2], cs->bc[3]}, {cs->period_i[0], cs->period_i[1]},
                  {cs->period_j[0], cs->period_j[1]}};
    if (build_geometry(gin, g) != 0) return -1;
    const long long N = (long long)ni * nj;
    for (long long c = 0; c < N; ++c) {
        cells3[3 * c] = g.cvol[c];
        cells3[3 * c + 1] = g.ccx[c];
        cells3[3 * c + 2] = g.ccy[c];
    }
    for (size_t k = 0; k < g.X.size(); ++k) X3[k] = g.X[k];
    for (size_t k = 0; k < g.fi.size(); ++k) fi[k] = g.fi[k];
    for (size_t k = 0; k < g.fj.size(); ++k) fj[k] = g.fj[k];
    return 0;
}

}   // extern "C"
