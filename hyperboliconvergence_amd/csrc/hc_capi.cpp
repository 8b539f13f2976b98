// hc_capi.cpp -- the C-ABI of include/hcamd.h: context, device buffers, the
// per-step launch sequence of the stage kernels, HIP graph capture, norms.
#include "hcamd.h"

#include "hc_geom.hpp"
#include "hc_launch.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

using namespace hc;

__global__ void k_reset_max(unsigned long long* p) { *p = dbl_bits(DBL_MIN); }

__global__ void k_clear_u32(unsigned int* p) { *p = 0u; }

// partial sums of (R/vol)^2 and max |R/vol| per variable; one partial per work-group
__global__ void __launch_bounds__(256) k_norm_partials(const double* res, const double* vol, long long N, int nv,
                                                       double* part)
{
    __shared__ double ss[4][256];
    __shared__ double sm[4][256];
    const int tid = threadIdx.x;
    double s[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
    for (long long c = (long long)blockIdx.x * 256 + tid; c < N; c += (long long)gridDim.x * 256) {
        for (int v = 0; v < nv; ++v) {
            const double r = res[c * (nv + 1) + v] / vol[c];
            s[v] = s[v] + r * r;
            m[v] = fmax(m[v], fabs(r));
        }
    }
    for (int v = 0; v < 4; ++v) {
        ss[v][tid] = s[v];
        sm[v][tid] = m[v];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int v = 0; v < 4; ++v) {
                ss[v][tid] = ss[v][tid] + ss[v][tid + w];
                sm[v][tid] = fmax(sm[v][tid], sm[v][tid + w]);
            }
        __syncthreads();
    }
    if (tid == 0)
        for (int v = 0; v < 4; ++v) {
            part[blockIdx.x * 8 + v] = ss[v][0];
            part[blockIdx.x * 8 + 4 + v] = sm[v][0];
        }
}

__global__ void __launch_bounds__(256) k_norm_final(const double* part, int nparts, long long N, double* out)
{
    __shared__ double ss[4][256];
    __shared__ double sm[4][256];
    const int tid = threadIdx.x;
    double s[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
    for (int p = tid; p < nparts; p += 256)
        for (int v = 0; v < 4; ++v) {
            s[v] = s[v] + part[p * 8 + v];
            m[v] = fmax(m[v], part[p * 8 + 4 + v]);
        }
    for (int v = 0; v < 4; ++v) {
        ss[v][tid] = s[v];
        sm[v][tid] = m[v];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int v = 0; v < 4; ++v) {
                ss[v][tid] = ss[v][tid] + ss[v][tid + w];
                sm[v][tid] = fmax(sm[v][tid], sm[v][tid + w]);
            }
        __syncthreads();
    }
    if (tid == 0)
        for (int v = 0; v < 4; ++v) {
            out[v] = sqrt(ss[v][0] / (double)N);
            out[4 + v] = sm[v][0];
        }
}

constexpr int NORM_BLOCKS = 1024;

__global__ void __launch_bounds__(256) k_triad(double* __restrict__ a, const double* __restrict__ b,
                                               const double* __restrict__ c, double s, long long n)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        a[i] = b[i] + s * c[i];
}

__global__ void __launch_bounds__(256) k_fill(double* a, double v, long long n)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) a[i] = v;
}

template <class T> void free_dev(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}   // namespace

struct hc_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // mesh and case
    int ni = 0, nj = 0;
    std::vector<double> nodes;
    bool have_mesh = false, have_case = false, have_state = false;
    hc_case_desc cs{};
    int nv = 0;
    int precision = HC_PRECISION_FP64;
    std::vector<RkPlan> plans;
    int nslots = 0;
    // device buffers
    double *cvol = nullptr, *ccx = nullptr, *ccy = nullptr, *fi = nullptr, *fj = nullptr;
    float *fi32 = nullptr, *fj32 = nullptr;
    double* q[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    double *acc = nullptr, *tmp = nullptr, *lam0 = nullptr, *bref = nullptr, *res = nullptr, *part = nullptr,
           *nout = nullptr;
    unsigned long long* maxslot = nullptr;
    double* dtslot = nullptr;
    unsigned int* status = nullptr;
    // graphs
    int graph_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_cur = -1;
    // profiling
    int profiling = 0;
    double prof[4] = {0, 0, 0, 0};
    std::vector<hipEvent_t> ev_pool;
    struct Pending {
        hipEvent_t e0, e1;
        double n_stage, n_update;   // launches the interval covers
        int kind;                   // 0 stage kernel, 1 update kernel, 2 whole hc_step call
    };
    std::vector<Pending> ev_pending;
};

namespace {

int fail(hc_context* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

int hipfail(hc_context* c, hipError_t e, const char* where)
{
    return fail(c, HC_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HC_HIP(call)                                           \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return hipfail(ctx, e_, #call);  \
    } while (0)

void release_buffers(hc_context* ctx)
{
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->gexec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_cur = -1;
    free_dev(ctx->cvol);
    free_dev(ctx->ccx);
    free_dev(ctx->ccy);
    free_dev(ctx->fi);
    free_dev(ctx->fj);
    free_dev(ctx->fi32);
    free_dev(ctx->fj32);
    for (auto& b : ctx->q) free_dev(b);
    free_dev(ctx->acc);
    free_dev(ctx->tmp);
    free_dev(ctx->lam0);
    free_dev(ctx->bref);
    free_dev(ctx->res);
}

bool valid_ids(const hc_case_desc& d)
{
    if (d.law < 0 || d.law > 1) return false;
    if (d.basis < 0 || d.basis > 1) return false;
    if (d.law == HC_LAW_SCALAR && (d.basis != HC_BASIS_CONSERVED || d.flux > HC_FLUX_RUSANOV)) return false;
    if (d.flux < 0 || d.flux > 5) return false;
    if (d.limiter < 0 || d.limiter > 9) return false;
    if (d.vscale < 0 || d.vscale > 2 || d.pscale < 0 || d.pscale > 2) return false;
    for (int b = 0; b < 4; ++b) {
        if (d.bc[b] < 0 || d.bc[b] > 4) return false;
        if (d.law == HC_LAW_SCALAR && (d.bc[b] == HC_BC_ENTROPY || d.bc[b] == HC_BC_WALL)) return false;
    }
    if ((d.bc[0] == HC_BC_PERIODIC) != (d.bc[1] == HC_BC_PERIODIC)) return false;
    if ((d.bc[2] == HC_BC_PERIODIC) != (d.bc[3] == HC_BC_PERIODIC)) return false;
    if (d.nstages < 1 || d.nstages > 6) return false;
    if (d.dt_mode < 0 || d.dt_mode > 1) return false;
    if (!(d.cfl > 0.0)) return false;
    return true;
}

void build_plans(hc_context* ctx)
{
    ctx->nslots = make_rk_plans(ctx->cs.nstages, ctx->cs.alpha, ctx->cs.beta, ctx->plans);
}

template <class RF> StageArgs<RF> base_args(hc_context* ctx)
{
    StageArgs<RF> a{};
    a.m.ni = ctx->ni;
    a.m.nj = ctx->nj;
    a.m.per_i = ctx->cs.bc[0] == HC_BC_PERIODIC;
    a.m.per_j = ctx->cs.bc[2] == HC_BC_PERIODIC;
    a.m.nfi = (long long)(ctx->ni + 1) * ctx->nj;
    a.m.nfj = (long long)ctx->ni * (ctx->nj + 1);
    a.m.cvol = ctx->cvol;
    a.m.ccx = ctx->ccx;
    a.m.ccy = ctx->ccy;
    if constexpr (sizeof(RF) == 8) {
        a.m.fi = ctx->fi;
        a.m.fj = ctx->fj;
    } else {
        a.m.fi = ctx->fi32;
        a.m.fj = ctx->fj32;
    }
    const hc_species& s = ctx->cs.sp;
    a.sp = Species<double>{s.gamma, s.minf, s.lref, s.nu, s.pr, s.dt, s.R, s.gamma1};
    a.spf = Species<RF>{RF(s.gamma), RF(s.minf), RF(s.lref), RF(s.nu), RF(s.pr), RF(s.dt), RF(s.R), RF(s.gamma1)};
    a.vscale = ctx->cs.vscale;
    a.pscale = ctx->cs.pscale;
    a.limiter = ctx->cs.limiter;
    for (int b = 0; b < 4; ++b) a.bc[b] = ctx->cs.bc[b];
    a.dt_mode = ctx->cs.dt_mode;
    a.N = (long long)ctx->ni * ctx->nj;
    a.cfl = ctx->cs.cfl;
    a.acc = ctx->acc;
    a.tmp = ctx->tmp;
    a.lam0 = ctx->lam0;
    a.bref = ctx->bref;
    a.res_out = ctx->res;
    a.maxslot = ctx->maxslot;
    a.dtslot = ctx->dtslot;
    a.status = ctx->status;
    return a;
}

hipError_t launch_stage(hc_context* ctx, const StageArgs<double>& a64, const StageArgs<float>& a32)
{
    const int key = ctx->cs.law * 100 + ctx->cs.basis * 10 + ctx->cs.flux;
    const bool mixed = ctx->precision == HC_PRECISION_MIXED;
#define HC_CASE_STAGE(L, B, F) \
    case L * 100 + B * 10 + F: \
        return mixed ? hc_stage_L##L##_B##B##_F##F##_P32(a32, ctx->stream) : hc_stage_L##L##_B##B##_F##F##_P64(a64, ctx->stream);
    switch (key) {
        HC_FOR_EACH_KERNEL(HC_CASE_STAGE)
    default: return hipErrorInvalidValue;
    }
#undef HC_CASE_STAGE
}

hipError_t launch_update(hc_context* ctx, const StageArgs<double>& a64, const StageArgs<float>& a32)
{
    const int key = ctx->cs.law * 10 + ctx->cs.basis;
    const bool mixed = ctx->precision == HC_PRECISION_MIXED;
#define HC_CASE_UPD(L, B) \
    case L * 10 + B: return mixed ? hc_upd_L##L##_B##B##_P32(a32, ctx->stream) : hc_upd_L##L##_B##B##_P64(a64, ctx->stream);
    switch (key) {
        HC_FOR_EACH_LAWBASIS(HC_CASE_UPD)
    default: return hipErrorInvalidValue;
    }
#undef HC_CASE_UPD
}

hipEvent_t take_event(hc_context* ctx)
{
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// profiling mode 1: events around every launch; kind 0 stage kernel, 1 update kernel
template <class F> hipError_t timed(hc_context* ctx, int kind, F&& launch)
{
    if (ctx->profiling != 1) return launch();
    hipEvent_t e0 = take_event(ctx), e1 = take_event(ctx);
    (void)hipEventRecord(e0, ctx->stream);
    const hipError_t e = launch();
    (void)hipEventRecord(e1, ctx->stream);
    ctx->ev_pending.push_back({e0, e1, kind == 0 ? 1.0 : 0.0, kind == 1 ? 1.0 : 0.0, kind});
    return e;
}

void drain_events(hc_context* ctx)
{
    for (auto& p : ctx->ev_pending) {
        float ms = 0.f;
        (void)hipEventSynchronize(p.e1);
        (void)hipEventElapsedTime(&ms, p.e0, p.e1);
        ctx->prof[p.kind == 1 ? 2 : 0] += ms;
        ctx->prof[1] += p.n_stage;
        ctx->prof[3] += p.n_update;
        ctx->ev_pool.push_back(p.e0);
        ctx->ev_pool.push_back(p.e1);
    }
    ctx->ev_pending.clear();
}

// the launches of one time step (timestepping/rungeKutta.h:97-164)
int enqueue_step(hc_context* ctx)
{
    const int ns = ctx->cs.nstages;
    const bool global = ctx->cs.dt_mode == HC_DT_GLOBAL;
    double* q0 = ctx->q[ctx->cur];
    double* bufA;
    double* bufB;
    if (ns == 1) {
        bufA = ctx->q[1 - ctx->cur];
        bufB = bufA;
    } else {
        bufA = ctx->q[(ctx->cur + 1) % 3];
        bufB = ctx->q[(ctx->cur + 2) % 3];
    }
    StageArgs<double> a64 = base_args<double>(ctx);
    StageArgs<float> a32 = base_args<float>(ctx);
    const double* in = q0;
    for (int s = 0; s < ns; ++s) {
        double* out;
        if (ns == 1) out = bufA;
        else if (s == ns - 1) out = q0;   // the last stage updates the step-start state in place
        else out = (s % 2 == 0) ? bufA : bufB;
        a64.q_in = in;
        a64.q0 = q0;
        a64.q_out = out;
        a64.stage = s;
        a64.plan = ctx->plans[s];
        a64.reset_max = (global && ns >= 2 && s == ns - 1) ? 1 : 0;
        a64.mode = (global && s == 0) ? MODE_GLOBAL_STAGE0 : MODE_FUSED;
        a32.q_in = a64.q_in;
        a32.q0 = a64.q0;
        a32.q_out = a64.q_out;
        a32.stage = s;
        a32.plan = a64.plan;
        a32.reset_max = a64.reset_max;
        a32.mode = a64.mode;
        hipError_t e = timed(ctx, 0, [&] { return launch_stage(ctx, a64, a32); });
        if (e != hipSuccess) return hipfail(ctx, e, "stage kernel launch");
        if (global && s == 0) {
            e = timed(ctx, 1, [&] { return launch_update(ctx, a64, a32); });
            if (e != hipSuccess) return hipfail(ctx, e, "update kernel launch");
        }
        in = out;
    }
    if (global && ns == 1) {
        hipLaunchKernelGGL(k_reset_max, dim3(1), dim3(1), 0, ctx->stream, ctx->maxslot);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hipfail(ctx, e, "reset kernel launch");
    }
    if (ns == 1) ctx->cur = 1 - ctx->cur;
    return HC_OK;
}

int ready(hc_context* ctx)
{
    if (!ctx) return HC_ERR_ARG;
    if (!ctx->have_mesh || !ctx->have_case) return fail(ctx, HC_ERR_STATE, "mesh and case must be set first");
    return HC_OK;
}

int check_status(hc_context* ctx)
{
    unsigned int st = 0;
    HC_HIP(hipMemcpyAsync(&st, ctx->status, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    if (st) return fail(ctx, HC_ERR_STATUS, "device status bits raised: " + std::to_string(st));
    return HC_OK;
}

}   // namespace

extern "C" {

const char* hc_version(void) { return "hyperboliconvergence_amd 0.1 (gfx950)"; }

int hc_create(hc_context** out, int device)
{
    if (!out) return HC_ERR_ARG;
    *out = nullptr;
    hc_context* ctx = new hc_context();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&ctx->maxslot, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&ctx->dtslot, 2 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->status, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMalloc(&ctx->part, NORM_BLOCKS * 8 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->nout, 8 * sizeof(double));
    if (e == hipSuccess) {
        const unsigned long long m = dbl_bits(DBL_MIN);
        const double z[2] = {0.0, 0.0};
        const unsigned int zs = 0;
        e = hipMemcpy(ctx->maxslot, &m, sizeof(m), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(ctx->dtslot, z, sizeof(z), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(ctx->status, &zs, sizeof(zs), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        hc_destroy(ctx);
        return HC_ERR_HIP;
    }
    *out = ctx;
    return HC_OK;
}

int hc_destroy(hc_context* ctx)
{
    if (!ctx) return HC_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    release_buffers(ctx);
    free_dev(ctx->maxslot);
    free_dev(ctx->dtslot);
    free_dev(ctx->status);
    free_dev(ctx->part);
    free_dev(ctx->nout);
    for (auto& p : ctx->ev_pending) {
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HC_OK;
}

const char* hc_last_error(const hc_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hc_set_mesh(hc_context* ctx, int ni, int nj, const double* nodes)
{
    if (!ctx || !nodes) return HC_ERR_ARG;
    if (ni < 3 || nj < 3) return fail(ctx, HC_ERR_ARG, "mesh needs at least 3x3 cells");
    if ((long long)ni * nj > (1LL << 31) - 1) return fail(ctx, HC_ERR_ARG, "mesh too large");
    HC_HIP(hipSetDevice(ctx->device));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    release_buffers(ctx);
    ctx->ni = ni;
    ctx->nj = nj;
    ctx->nodes.assign(nodes, nodes + 2 * (size_t)(ni + 1) * (nj + 1));
    ctx->have_mesh = true;
    ctx->have_case = false;
    ctx->have_state = false;
    return HC_OK;
}

int hc_set_case(hc_context* ctx, const hc_case_desc* desc)
{
    if (!ctx || !desc) return HC_ERR_ARG;
    if (!ctx->have_mesh) return fail(ctx, HC_ERR_STATE, "hc_set_mesh must come first");
    if (!valid_ids(*desc)) return fail(ctx, HC_ERR_ARG, "invalid case descriptor");
    HC_HIP(hipSetDevice(ctx->device));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    const int ni = ctx->ni, nj = ctx->nj;
    GeomInput gin{ni, nj, ctx->nodes.data(), {desc->bc[0], desc->bc[1], desc->bc[2], desc->bc[3]},
                  {desc->period_i[0], desc->period_i[1]}, {desc->period_j[0], desc->period_j[1]}};
    GeomOutput g;
    if (build_geometry(gin, g) != 0) return fail(ctx, HC_ERR_ARG, "degenerate mesh cell");
    // keep the state across a case change when the variable count is unchanged
    std::vector<double> keep;
    const int nv_new = desc->law == HC_LAW_EULER ? 4 : 3;
    const long long N = (long long)ni * nj;
    if (ctx->have_state && ctx->nv == nv_new) {
        keep.resize((size_t)nv_new * N);
        HC_HIP(hipMemcpy(keep.data(), ctx->q[ctx->cur], keep.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    release_buffers(ctx);
    ctx->cs = *desc;
    ctx->nv = nv_new;
    build_plans(ctx);
    const long long nfi = (long long)(ni + 1) * nj, nfj = (long long)ni * (nj + 1);
    HC_HIP(hipMalloc(&ctx->cvol, N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->ccx, N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->ccy, N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->fi, NSLOT * nfi * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->fj, NSLOT * nfj * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->fi32, NSLOT * nfi * sizeof(float)));
    HC_HIP(hipMalloc(&ctx->fj32, NSLOT * nfj * sizeof(float)));
    HC_HIP(hipMemcpy(ctx->cvol, g.cvol.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HC_HIP(hipMemcpy(ctx->ccx, g.ccx.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HC_HIP(hipMemcpy(ctx->ccy, g.ccy.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HC_HIP(hipMemcpy(ctx->fi, g.fi.data(), g.fi.size() * sizeof(double), hipMemcpyHostToDevice));
    HC_HIP(hipMemcpy(ctx->fj, g.fj.data(), g.fj.size() * sizeof(double), hipMemcpyHostToDevice));
    {
        std::vector<float> f32(g.fi.begin(), g.fi.end());
        HC_HIP(hipMemcpy(ctx->fi32, f32.data(), f32.size() * sizeof(float), hipMemcpyHostToDevice));
        f32.assign(g.fj.begin(), g.fj.end());
        HC_HIP(hipMemcpy(ctx->fj32, f32.data(), f32.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    const int nv = ctx->nv;
    for (auto& b : ctx->q) HC_HIP(hipMalloc(&b, (size_t)nv * N * sizeof(double)));
    ctx->cur = 0;
    if (ctx->nslots > 0) HC_HIP(hipMalloc(&ctx->acc, (size_t)ctx->nslots * nv * N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->tmp, (size_t)nv * N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->lam0, N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->res, (size_t)(nv + 1) * N * sizeof(double)));
    const size_t nb = (size_t)(2 * nj + 2 * ni) * nv;
    HC_HIP(hipMalloc(&ctx->bref, nb * sizeof(double)));
    std::vector<double> zeros(nb, 0.0);
    HC_HIP(hipMemcpy(ctx->bref, zeros.data(), nb * sizeof(double), hipMemcpyHostToDevice));
    if (!keep.empty()) {
        HC_HIP(hipMemcpy(ctx->q[0], keep.data(), keep.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        ctx->have_state = false;
    }
    const unsigned long long m = dbl_bits(DBL_MIN);
    HC_HIP(hipMemcpy(ctx->maxslot, &m, sizeof(m), hipMemcpyHostToDevice));
    ctx->have_case = true;
    return HC_OK;
}

int hc_set_precision(hc_context* ctx, int precision)
{
    if (!ctx || (precision != HC_PRECISION_FP64 && precision != HC_PRECISION_MIXED)) return HC_ERR_ARG;
    if (ctx->precision != precision) {
        if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
        if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
        ctx->gexec = nullptr;
        ctx->graph = nullptr;
        ctx->graph_cur = -1;
    }
    ctx->precision = precision;
    return HC_OK;
}

int hc_set_boundary_refs(hc_context* ctx, const double* bref)
{
    int r = ready(ctx);
    if (r) return r;
    if (!bref) return fail(ctx, HC_ERR_ARG, "null boundary reference array");
    const size_t nb = (size_t)(2 * ctx->nj + 2 * ctx->ni) * ctx->nv;
    HC_HIP(hipMemcpyAsync(ctx->bref, bref, nb * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_set_state(hc_context* ctx, const double* q)
{
    int r = ready(ctx);
    if (r) return r;
    if (!q) return fail(ctx, HC_ERR_ARG, "null state");
    const long long N = (long long)ctx->ni * ctx->nj;
    const int nv = ctx->nv;
    std::vector<double> soa((size_t)nv * N);
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < nv; ++v) soa[(size_t)v * N + c] = q[c * nv + v];
    HC_HIP(hipStreamSynchronize(ctx->stream));
    HC_HIP(hipMemcpy(ctx->q[ctx->cur], soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
    const double z[2] = {0.0, 0.0};
    HC_HIP(hipMemcpy(ctx->dtslot, z, sizeof(z), hipMemcpyHostToDevice));
    ctx->have_state = true;
    return HC_OK;
}

int hc_get_state(hc_context* ctx, double* q)
{
    int r = ready(ctx);
    if (r) return r;
    if (!q || !ctx->have_state) return fail(ctx, HC_ERR_STATE, "no state");
    const long long N = (long long)ctx->ni * ctx->nj;
    const int nv = ctx->nv;
    std::vector<double> soa((size_t)nv * N);
    HC_HIP(hipStreamSynchronize(ctx->stream));
    HC_HIP(hipMemcpy(soa.data(), ctx->q[ctx->cur], soa.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < nv; ++v) q[c * nv + v] = soa[(size_t)v * N + c];
    return HC_OK;
}

static int residual_to_device(hc_context* ctx)
{
    StageArgs<double> a64 = base_args<double>(ctx);
    StageArgs<float> a32 = base_args<float>(ctx);
    a64.mode = a32.mode = MODE_RESIDUAL;
    a64.q_in = a32.q_in = ctx->q[ctx->cur];
    a64.q0 = a32.q0 = ctx->q[ctx->cur];
    a64.plan = a32.plan = ctx->plans[0];
    const hipError_t e = launch_stage(ctx, a64, a32);
    if (e != hipSuccess) return hipfail(ctx, e, "residual kernel launch");
    return HC_OK;
}

int hc_residual(hc_context* ctx, double* res)
{
    int r = ready(ctx);
    if (r) return r;
    if (!res || !ctx->have_state) return fail(ctx, HC_ERR_STATE, "no state");
    r = residual_to_device(ctx);
    if (r) return r;
    const long long N = (long long)ctx->ni * ctx->nj;
    HC_HIP(hipMemcpyAsync(res, ctx->res, (size_t)(ctx->nv + 1) * N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return check_status(ctx);
}

int hc_norms(hc_context* ctx, double* l2, double* linf)
{
    int r = ready(ctx);
    if (r) return r;
    if (!l2 || !linf || !ctx->have_state) return fail(ctx, HC_ERR_STATE, "no state");
    r = residual_to_device(ctx);
    if (r) return r;
    const long long N = (long long)ctx->ni * ctx->nj;
    hipLaunchKernelGGL(k_norm_partials, dim3(NORM_BLOCKS), dim3(256), 0, ctx->stream, ctx->res, ctx->cvol, N, ctx->nv,
                       ctx->part);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, ctx->stream, ctx->part, NORM_BLOCKS, N, ctx->nout);
    HC_HIP(hipGetLastError());
    double out[8];
    HC_HIP(hipMemcpyAsync(out, ctx->nout, sizeof(out), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    for (int v = 0; v < ctx->nv; ++v) {
        l2[v] = out[v];
        linf[v] = out[4 + v];
    }
    return check_status(ctx);
}

int hc_step(hc_context* ctx, int nsteps)
{
    int r = ready(ctx);
    if (r) return r;
    if (nsteps < 0) return fail(ctx, HC_ERR_ARG, "negative step count");
    if (!ctx->have_state) return fail(ctx, HC_ERR_STATE, "no state");
    int done = 0;
    int k = ctx->graph_steps;
    // single-stage tables alternate two state buffers: capture an even number of steps
    if (ctx->cs.nstages == 1 && (k % 2)) k += 1;
    hipEvent_t r0 = nullptr;
    if (ctx->profiling == 2 && nsteps > 0) {
        r0 = take_event(ctx);
        (void)hipEventRecord(r0, ctx->stream);
    }
    if (k > 0 && ctx->profiling != 1) {
        if (ctx->gexec && ctx->graph_cur != ctx->cur) {
            (void)hipGraphExecDestroy(ctx->gexec);
            (void)hipGraphDestroy(ctx->graph);
            ctx->gexec = nullptr;
            ctx->graph = nullptr;
        }
        while (nsteps - done >= k) {
            if (!ctx->gexec) {
                const int cur0 = ctx->cur;
                HC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                int rc = HC_OK;
                for (int s = 0; s < k && rc == HC_OK; ++s) rc = enqueue_step(ctx);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
                ctx->cur = cur0;
                if (rc != HC_OK) return rc;
                if (ec != hipSuccess) return hipfail(ctx, ec, "hipStreamEndCapture");
                ctx->graph = g;
                HC_HIP(hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0));
                ctx->graph_cur = cur0;
            }
            HC_HIP(hipGraphLaunch(ctx->gexec, ctx->stream));
            done += k;
        }
    }
    for (; done < nsteps; ++done) {
        r = enqueue_step(ctx);
        if (r) return r;
    }
    if (r0) {
        hipEvent_t r1 = take_event(ctx);
        (void)hipEventRecord(r1, ctx->stream);
        const double upd = ctx->cs.dt_mode == 0 ? 1.0 : 0.0;
        ctx->ev_pending.push_back({r0, r1, double(nsteps) * ctx->cs.nstages, double(nsteps) * upd, 2});
    }
    if (ctx->profiling == 1) drain_events(ctx);
    return HC_OK;
}

int hc_get_time(hc_context* ctx, double* t)
{
    int r = ready(ctx);
    if (r) return r;
    if (!t) return HC_ERR_ARG;
    double d[2];
    HC_HIP(hipMemcpyAsync(d, ctx->dtslot, sizeof(d), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    *t = d[1];
    return HC_OK;
}

int hc_get_status(hc_context* ctx, unsigned int* status)
{
    if (!ctx || !status) return HC_ERR_ARG;
    HC_HIP(hipMemcpyAsync(status, ctx->status, sizeof(unsigned int), hipMemcpyDeviceToHost, ctx->stream));
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(1), 0, ctx->stream, ctx->status);
    HC_HIP(hipGetLastError());
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_synchronize(hc_context* ctx)
{
    if (!ctx) return HC_ERR_ARG;
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_set_graph_steps(hc_context* ctx, int k)
{
    if (!ctx || k < 0) return HC_ERR_ARG;
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->gexec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_cur = -1;
    ctx->graph_steps = k;
    return HC_OK;
}

int hc_set_profiling(hc_context* ctx, int enable)
{
    if (!ctx) return HC_ERR_ARG;
    if (enable < 0 || enable > 2) return fail(ctx, HC_ERR_ARG, "profiling mode must be 0, 1 or 2");
    ctx->profiling = enable;
    for (double& p : ctx->prof) p = 0.0;
    return HC_OK;
}

int hc_get_profile(hc_context* ctx, double* out4)
{
    if (!ctx || !out4) return HC_ERR_ARG;
    drain_events(ctx);
    for (int k = 0; k < 4; ++k) out4[k] = ctx->prof[k];
    return HC_OK;
}

int hc_stream_triad(hc_context* ctx, long long n, int reps, double* gbps)
{
    if (!ctx || n <= 0 || reps <= 0 || !gbps) return HC_ERR_ARG;
    HC_HIP(hipSetDevice(ctx->device));
    double *a = nullptr, *b = nullptr, *c = nullptr;
    HC_HIP(hipMalloc(&a, n * sizeof(double)));
    HC_HIP(hipMalloc(&b, n * sizeof(double)));
    HC_HIP(hipMalloc(&c, n * sizeof(double)));
    const unsigned grid = 256 * 32;
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, ctx->stream, b, 1.0, n);
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, ctx->stream, c, 2.0, n);
    hipLaunchKernelGGL(k_triad, dim3(grid), dim3(256), 0, ctx->stream, a, b, c, 3.0, n);   // warm-up
    hipEvent_t e0, e1;
    HC_HIP(hipEventCreate(&e0));
    HC_HIP(hipEventCreate(&e1));
    HC_HIP(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_triad, dim3(grid), dim3(256), 0, ctx->stream, a, b, c, 3.0, n);
    HC_HIP(hipEventRecord(e1, ctx->stream));
    HC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    HC_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(c);
    *gbps = 24.0 * (double)n * reps / (ms * 1e-3) / 1e9;
    return HC_OK;
}

int hc_device_state(hc_context* ctx, double** dptr)
{
    int r = ready(ctx);
    if (r) return r;
    if (!dptr) return HC_ERR_ARG;
    *dptr = ctx->q[ctx->cur];
    return HC_OK;
}

}   // extern "C"
