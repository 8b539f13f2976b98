// hc_adjoint.cpp -- tangent and adjoint entry points of the C-ABI (include/hcamd.h):
// residual-level products with the Jacobian and its transpose, and the tangent / adjoint of
// K explicit Runge-Kutta steps with a fixed time step in the conserved basis, the latter
// with checkpoints every C steps (sqrt(K) memory: the states of one segment are recomputed
// from its checkpoint on the way back).
//
// Time stepping (timestepping/rungeKutta.h:97-164 with a fixed dt), per stage s:
//   q_{s+1} = q^n + (beta_s dt / V) sum_{k<=s} alpha[s][k] R(q_k),  q_0 = q^n,  q^{n+1} = q_ns.
// Reverse sweep of one step, a_s = adjoint of q_{s+1}, ybar_k = adjoint of R(q_k):
//   for s = ns-1 .. 0:  qbar^n += a_s;  ybar_k += alpha[s][k] (beta_s dt / V) a_s  (k <= s);
//                       g = (dR/dq(q_s))^T ybar_s;  a_{s-1} += g (s > 0),  qbar^n += g (s = 0)
// and the parameter adjoints (minf, boundary references) collect (dR/dp)^T ybar_s.
#include "hc_adjoint_launch.hpp"
#include "hc_context.hpp"

#include <cstring>
#include <utility>

using namespace hc;
using namespace hc_detail;

struct AdjState {
    long long N = 0, nfaces = 0;
    int nv = 0, ns = 0;
    // residual-level work arrays
    double *Q = nullptr, *Qd = nullptr, *F = nullptr, *Fd = nullptr, *res = nullptr, *resd = nullptr;
    double *minf_face = nullptr, *qbar = nullptr, *Qbar = nullptr;
    double *ybar_in = nullptr, *qd_in = nullptr, *brefd_in = nullptr, *part = nullptr, *scal = nullptr;
    // trajectory arrays
    std::vector<double*> ckpt;       // checkpoints, one state each
    std::vector<double*> seg;        // stage states of one segment, C x ns
    std::vector<double*> stage;      // stage states q_1 .. q_{ns-1} (tangent steps)
    std::vector<double*> a;          // the second step-boundary adjoint of the backward sweep
    std::vector<double*> yb;         // adjoints of the stage residuals
    std::vector<double*> rd;         // tangent residuals of the stages (AoS [N][nv])
    std::vector<double*> qdot;       // stage tangents q_dot_0 .. q_dot_{ns-1}
    double *qbn = nullptr, *brefbar_tot = nullptr, *qtmp = nullptr;
    int K = 0, C = 0;                // recorded trajectory
    bool have_traj = false;
};

namespace {

constexpr int SUM_BLOCKS = 256;

__global__ void __launch_bounds__(256) k_sum_partials(const double* x, long long n, double* part)
{
    __shared__ double s[256];
    double a = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) a = a + x[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// out[slot] (+)= sum of the partials (one work-group, fixed order)
__global__ void __launch_bounds__(256) k_sum_final(const double* part, int n, double* out, int slot, int accumulate)
{
    __shared__ double s[256];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) a = a + part[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot] = accumulate ? out[slot] + s[0] : s[0];
}

// Runge-Kutta adjoint of one stage (rk_deliver) over every unknown
__global__ void __launch_bounds__(256) k_rk_adjoint(const double* x, RkDeliver t, const double* vol, long long N, int nv)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < N * nv) rk_deliver(t, i, x[i], vol[i % N]);
}

// stage tangent: out[v][c] = base[v][c] + (beta dt / V_c) sum_k coef[k] rd_k[c][v]
struct RkTan {
    const double* rd[6];
    double coef[6];
    int nk;
};
__global__ void __launch_bounds__(256) k_rk_tangent(double* out, const double* base, RkTan t, double bdt,
                                                    const double* vol, long long N, int nv)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= N * nv) return;
    const long long c = i % N;
    const int v = (int)(i / N);
    double s = 0.0;
    for (int k = 0; k < t.nk; ++k) s = s + t.coef[k] * t.rd[k][c * nv + v];
    out[i] = base[i] + (bdt / vol[c]) * s;
}

unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

template <class T> void free_d(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int adj_ready(hc_context* ctx, bool need_traj_basis)
{
    int r = case_ready(ctx, 0);
    if (r) return r;
    if (ctx->cases.size() != 1) return fail(ctx, HC_ERR_STATE, "tangent / adjoint entry points take a single-case context");
    if (!ctx->cases[0].have_state) return fail(ctx, HC_ERR_STATE, "no state");
    const hc_case_desc& d = ctx->cs;
    const int key = d.law * 100 + d.basis * 10 + d.flux;
#define HC_KEY_OK(L, B, F) key == L * 100 + B * 10 + F ||
    if (!(HC_FOR_EACH_ADJ(HC_KEY_OK) false))
        return fail(ctx, HC_ERR_ARG, "no tangent / adjoint kernels for this law, basis and flux");
#undef HC_KEY_OK
    if (need_traj_basis && (d.basis != HC_BASIS_CONSERVED || d.dt_mode != HC_DT_FIXED))
        return fail(ctx, HC_ERR_ARG, "trajectory derivatives need the conserved basis and a fixed time step");
    if (ctx->precision != HC_PRECISION_FP64) return fail(ctx, HC_ERR_ARG, "tangent / adjoint run in FP64");
    return HC_OK;
}

int alloc(hc_context* ctx, double*& p, long long n)
{
    if (p) return HC_OK;
    HC_HIP(hipMalloc(&p, (size_t)(n > 0 ? n : 1) * sizeof(double)));
    HC_HIP(hipMemset(p, 0, (size_t)(n > 0 ? n : 1) * sizeof(double)));
    return HC_OK;
}

int ensure(hc_context* ctx)
{
    if (!ctx->adj) ctx->adj = new AdjState();
    AdjState& s = *ctx->adj;
    const auto& c = ctx->cases[0];
    s.N = c.N;
    s.nfaces = c.nfi + c.nfj;
    s.nv = ctx->nv;
    s.ns = ctx->cs.nstages;
    const long long N = s.N, nv = s.nv, nF = s.nfaces;
    int r = 0;
    if ((r = alloc(ctx, s.Q, 2 * nv * N)) || (r = alloc(ctx, s.Qd, 2 * nv * N)) || (r = alloc(ctx, s.F, (nv + 1) * nF)) ||
        (r = alloc(ctx, s.Fd, nv * nF)) || (r = alloc(ctx, s.res, (nv + 1) * N)) || (r = alloc(ctx, s.resd, nv * N)) ||
        (r = alloc(ctx, s.minf_face, nF)) ||
        (r = alloc(ctx, s.qbar, nv * N)) ||
        (r = alloc(ctx, s.Qbar, 2 * nv * N)) || (r = alloc(ctx, s.ybar_in, nv * N)) || (r = alloc(ctx, s.qd_in, nv * N)) ||
        (r = alloc(ctx, s.brefd_in, (long long)c.brows * nv)) || (r = alloc(ctx, s.part, SUM_BLOCKS)) ||
        (r = alloc(ctx, s.scal, 4)) || (r = alloc(ctx, s.qbn, nv * N)) ||
        (r = alloc(ctx, s.brefbar_tot, (long long)c.brows * nv)) || (r = alloc(ctx, s.qtmp, nv * N)))
        return r;
    return HC_OK;
}

AdjArgs base_adj(hc_context* ctx)
{
    const auto& c = ctx->cases[0];
    AdjState& s = *ctx->adj;
    AdjArgs a{};
    a.m.ni = c.ni;
    a.m.nj = c.nj;
    a.m.per_i = c.cs.bc[0] == HC_BC_PERIODIC;
    a.m.per_j = c.cs.bc[2] == HC_BC_PERIODIC;
    a.m.nfi = c.nfi;
    a.m.nfj = c.nfj;
    a.m.cvol = ctx->cvol;
    a.m.ccx = ctx->ccx;
    a.m.ccy = ctx->ccy;
    a.m.fi = ctx->fi;
    a.m.fj = ctx->fj;
    a.m.fi64 = ctx->fi;
    a.m.fj64 = ctx->fj;
    const hc_species& sp = c.cs.sp;
    a.sp = Species<double>{sp.gamma, sp.minf, sp.lref, sp.nu, sp.pr, sp.dt, sp.R, sp.gamma1};
    a.vscale = c.cs.vscale;
    a.pscale = c.cs.pscale;
    a.limiter = c.cs.limiter;
    for (int b = 0; b < 4; ++b) a.bc[b] = c.cs.bc[b];
    a.N = s.N;
    a.nfaces = s.nfaces;
    a.bref = ctx->bref;
    a.Q = s.Q;
    a.Qd = s.Qd;
    a.F = s.F;
    a.Fd = s.Fd;
    a.res = s.res;
    a.resd = s.resd;
    a.minf_face = s.minf_face;
    a.brefbar = s.brefbar_tot;
    a.qbar = s.qbar;
    a.Qbar = s.Qbar;
    return a;
}

hipError_t run_tangent(hc_context* ctx, const AdjArgs& a)
{
    const int key = ctx->cs.law * 100 + ctx->cs.basis * 10 + ctx->cs.flux;
#define HC_CASE_TAN(L, B, F) \
    case L * 100 + B * 10 + F: return hc_tan_L##L##_B##B##_F##F(a, ctx->stream);
    switch (key) {
        HC_FOR_EACH_ADJ(HC_CASE_TAN)
    default: return hipErrorInvalidValue;
    }
#undef HC_CASE_TAN
}

hipError_t run_adjoint(hc_context* ctx, const AdjArgs& a)
{
    const int key = ctx->cs.law * 100 + ctx->cs.basis * 10 + ctx->cs.flux;
#define HC_CASE_ADJ(L, B, F) \
    case L * 100 + B * 10 + F: return hc_adj_L##L##_B##B##_F##F(a, ctx->stream);
    switch (key) {
        HC_FOR_EACH_ADJ(HC_CASE_ADJ)
    default: return hipErrorInvalidValue;
    }
#undef HC_CASE_ADJ
}

// AoS [N][nv] host <-> SoA [nv][N] device
int upload_soa(hc_context* ctx, double* dst, const double* aos, long long N, int nv)
{
    std::vector<double> soa((size_t)nv * N);
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < nv; ++v) soa[(size_t)v * N + c] = aos[c * nv + v];
    HC_HIP(hipStreamSynchronize(ctx->stream));
    HC_HIP(hipMemcpy(dst, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
    return HC_OK;
}

int download_soa(hc_context* ctx, double* aos, const double* src, long long N, int nv)
{
    std::vector<double> soa((size_t)nv * N);
    HC_HIP(hipStreamSynchronize(ctx->stream));
    HC_HIP(hipMemcpy(soa.data(), src, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (long long c = 0; c < N; ++c)
        for (int v = 0; v < nv; ++v) aos[c * nv + v] = soa[(size_t)v * N + c];
    return HC_OK;
}

// adjoint residual at state q (device SoA) with ybar (device SoA): qbar delivered as the Runge-Kutta
// adjoint rk when given, else into dst (assigned, or added when add is set); the face minf adjoints
// added to s.minf_face, the boundary-reference adjoint to s.brefbar_tot
int adjoint_at(hc_context* ctx, const double* q, const double* ybar, double* dst, bool add, const RkDeliver* rk)
{
    AdjState& s = *ctx->adj;
    AdjArgs a = base_adj(ctx);
    a.q = q;
    a.ybar = ybar;
    a.brefbar = s.brefbar_tot;
    a.bref_acc = 1;
    a.qacc = dst;
    a.qacc_add = add ? 1 : 0;
    if (rk) {
        a.rk_on = 1;
        a.rk = *rk;
    }
    HC_HIP(run_adjoint(ctx, a));
    return HC_OK;
}

// minf_bar = the sum of the per-face minf adjoints, into s.scal[0]
int minf_total(hc_context* ctx)
{
    AdjState& s = *ctx->adj;
    hipLaunchKernelGGL(k_sum_partials, dim3(SUM_BLOCKS), dim3(256), 0, ctx->stream, s.minf_face, s.nfaces, s.part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, ctx->stream, s.part, SUM_BLOCKS, s.scal, 0, 0);
    HC_HIP(hipGetLastError());
    return HC_OK;
}

// stage s of the fused kernel from q_in into q_out (not in place), plans and accumulators as in a step
int stage_into(hc_context* ctx, int s, const double* q_in, const double* q0, double* q_out)
{
    LaunchArgs<double> L = launch_args<double>(ctx);
    L.q_in = q_in;
    L.q0 = q0;
    L.q_out = q_out;
    L.stage = s;
    L.plan = ctx->plans[s];
    L.reset_max = 0;
    L.mode = MODE_FUSED;
    const hipError_t e = launch_stage(ctx, L);
    if (e != hipSuccess) return hipfail(ctx, e, "stage kernel launch");
    return HC_OK;
}

int copy_d(hc_context* ctx, double* dst, const double* src, long long n)
{
    HC_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return HC_OK;
}

int grow(hc_context* ctx, std::vector<double*>& v, size_t n, long long len)
{
    while (v.size() < n) {
        double* p = nullptr;
        HC_HIP(hipMalloc(&p, (size_t)len * sizeof(double)));
        v.push_back(p);
    }
    return HC_OK;
}

}   // namespace

namespace hc_detail {

void release_adjoint(hc_context* ctx)
{
    if (!ctx->adj) return;
    AdjState& s = *ctx->adj;
    for (double** p : {&s.Q, &s.Qd, &s.F, &s.Fd, &s.res, &s.resd, &s.minf_face, &s.qbar, &s.Qbar,
                       &s.ybar_in, &s.qd_in, &s.brefd_in, &s.part, &s.scal, &s.qbn, &s.brefbar_tot, &s.qtmp})
        free_d(*p);
    for (auto* v : {&s.ckpt, &s.seg, &s.stage, &s.a, &s.yb, &s.rd, &s.qdot})
        for (double*& p : *v) free_d(p);
    delete ctx->adj;
    ctx->adj = nullptr;
}

}   // namespace hc_detail

extern "C" {

int hc_tangent_residual(hc_context* ctx, const double* qdot, const double* bref_dot, double minf_dot, double* res,
                        double* res_dot)
{
    int r = adj_ready(ctx, false);
    if (r) return r;
    if (!qdot || !res || !res_dot) return fail(ctx, HC_ERR_ARG, "null argument");
    if ((r = ensure(ctx))) return r;
    AdjState& s = *ctx->adj;
    const auto& c = ctx->cases[0];
    if ((r = upload_soa(ctx, s.qd_in, qdot, s.N, s.nv))) return r;
    if (bref_dot)
        HC_HIP(hipMemcpy(s.brefd_in, bref_dot, (size_t)c.brows * s.nv * sizeof(double), hipMemcpyHostToDevice));
    AdjArgs a = base_adj(ctx);
    a.q = ctx->q[ctx->cur];
    a.qd = s.qd_in;
    a.brefd = bref_dot ? s.brefd_in : nullptr;
    a.minfd = minf_dot;
    HC_HIP(run_tangent(ctx, a));
    HC_HIP(hipMemcpyAsync(res, s.res, (size_t)(s.nv + 1) * s.N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipMemcpyAsync(res_dot, s.resd, (size_t)s.nv * s.N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_adjoint_residual(hc_context* ctx, const double* ybar, double* qbar, double* bref_bar, double* minf_bar)
{
    int r = adj_ready(ctx, false);
    if (r) return r;
    if (!ybar || !qbar || !bref_bar || !minf_bar) return fail(ctx, HC_ERR_ARG, "null argument");
    if ((r = ensure(ctx))) return r;
    AdjState& s = *ctx->adj;
    const auto& c = ctx->cases[0];
    if ((r = upload_soa(ctx, s.ybar_in, ybar, s.N, s.nv))) return r;
    const long long nb = (long long)c.brows * s.nv;
    HC_HIP(hipMemsetAsync(s.brefbar_tot, 0, nb * sizeof(double), ctx->stream));
    HC_HIP(hipMemsetAsync(s.minf_face, 0, s.nfaces * sizeof(double), ctx->stream));
    if ((r = adjoint_at(ctx, ctx->q[ctx->cur], s.ybar_in, s.qbar, false, nullptr))) return r;
    if ((r = minf_total(ctx))) return r;
    if ((r = download_soa(ctx, qbar, s.qbar, s.N, s.nv))) return r;
    HC_HIP(hipMemcpy(bref_bar, s.brefbar_tot, nb * sizeof(double), hipMemcpyDeviceToHost));
    HC_HIP(hipMemcpy(minf_bar, s.scal, sizeof(double), hipMemcpyDeviceToHost));
    return HC_OK;
}

int hc_adjoint_forward(hc_context* ctx, int nsteps, int checkpoint_every)
{
    int r = adj_ready(ctx, true);
    if (r) return r;
    if (nsteps < 1 || checkpoint_every < 1) return fail(ctx, HC_ERR_ARG, "need nsteps >= 1 and checkpoint_every >= 1");
    if ((r = ensure(ctx))) return r;
    AdjState& s = *ctx->adj;
    const long long len = (long long)s.nv * s.N;
    const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every;
    if ((r = grow(ctx, s.ckpt, nck, len))) return r;
    for (int k = 0; k < nck; ++k) {
        if ((r = copy_d(ctx, s.ckpt[k], ctx->q[ctx->cur], len))) return r;
        const int n = std::min(checkpoint_every, nsteps - k * checkpoint_every);
        for (int t = 0; t < n; ++t)
            if ((r = enqueue_step(ctx))) return r;
    }
    HC_HIP(hipStreamSynchronize(ctx->stream));
    s.K = nsteps;
    s.C = checkpoint_every;
    s.have_traj = true;
    return check_status(ctx);
}

// dev: qbar_final, qbar_initial and bref_bar are device arrays (minf_bar stays a host scalar)
static int adjoint_backward(hc_context* ctx, const double* qbar_final, double* qbar_initial, double* bref_bar,
                            double* minf_bar, bool dev)
{
    int r = adj_ready(ctx, true);
    if (r) return r;
    if (!ctx->adj || !ctx->adj->have_traj) return fail(ctx, HC_ERR_STATE, "hc_adjoint_forward must come first");
    if (!qbar_final || !qbar_initial || !bref_bar || !minf_bar) return fail(ctx, HC_ERR_ARG, "null argument");
    if (dev && !(device_pointer(qbar_final) && device_pointer(qbar_initial) && device_pointer(bref_bar)))
        return fail(ctx, HC_ERR_ARG, "qbar_final, qbar_initial and bref_bar must be device memory of the library's HIP runtime");
    AdjState& s = *ctx->adj;
    const auto& c = ctx->cases[0];
    const int ns = s.ns, nv = s.nv;
    const long long N = s.N, len = (long long)nv * N, nb = (long long)c.brows * nv;
    const double dt = ctx->cs.cfl;
    // every stage state of a segment is kept (C steps x ns stages), so the reverse sweep recomputes
    // each step once, from the segment's checkpoint
    if ((r = grow(ctx, s.seg, (size_t)s.C * ns, len)) || (r = grow(ctx, s.a, 1, len)) || (r = grow(ctx, s.yb, ns, len)))
        return r;
    // lam holds the adjoint of the state at the current step boundary; nxt receives the one before
    double* lam = s.qbn;
    double* nxt = s.a[0];
    if (dev) {
        HC_HIP(hipStreamSynchronize(ctx->stream));
        if ((r = aos_to_soa_device(ctx, lam, qbar_final, N, nv))) return r;
    } else if ((r = upload_soa(ctx, lam, qbar_final, N, nv))) {
        return r;
    }
    HC_HIP(hipMemsetAsync(s.brefbar_tot, 0, nb * sizeof(double), ctx->stream));
    HC_HIP(hipMemsetAsync(s.minf_face, 0, s.nfaces * sizeof(double), ctx->stream));
    const int nseg = (s.K + s.C - 1) / s.C;
    for (int g = nseg - 1; g >= 0; --g) {
        const int n0 = g * s.C;
        const int n = std::min(s.C, s.K - n0);
        // the stage states of steps n0 .. n0+n-1 from the segment's checkpoint: seg[t ns + st] is the
        // input of stage st of step t (st = 0: the step's start state)
        if ((r = copy_d(ctx, s.seg[0], s.ckpt[g], len))) return r;
        for (int t = 0; t < n; ++t) {
            const double* q0 = s.seg[(size_t)t * ns];
            const double* in = q0;
            for (int st = 0; st < ns; ++st) {
                if (st == ns - 1 && t + 1 == n) break;   // the next segment's checkpoint is not needed
                double* out = st == ns - 1 ? s.seg[(size_t)(t + 1) * ns] : s.seg[(size_t)t * ns + st + 1];
                if ((r = stage_into(ctx, st, in, q0, out))) return r;
                in = out;
            }
        }
        for (int t = n - 1; t >= 0; --t) {
            // Runge-Kutta adjoint of stage st: its output's adjoint goes to nxt (the step's start
            // state enters every stage) and, scaled, to the residual adjoints yb[k] it combines
            bool yb_set[6] = {false, false, false, false, false, false};
            auto program = [&](int st) {
                RkDeliver d{};
                d.lam = nxt;
                d.lam_add = st < ns - 1 ? 1 : 0;
                for (int k = 0; k <= st; ++k) {
                    const double alpha = ctx->cs.alpha[st][k];
                    if (alpha == 0.0) continue;
                    d.yb[d.nk] = s.yb[k];
                    d.coef[d.nk] = alpha * (ctx->cs.beta[st] * dt);
                    d.add[d.nk] = yb_set[k] ? 1 : 0;
                    d.nk++;
                    yb_set[k] = true;
                }
                return d;
            };
            // the last stage's output is the step's end state, whose adjoint is lam
            const RkDeliver last = program(ns - 1);
            hipLaunchKernelGGL(k_rk_adjoint, dim3(blocks(len)), dim3(256), 0, ctx->stream, lam, last, ctx->cvol, N, nv);
            HC_HIP(hipGetLastError());
            for (int st = ns - 1; st >= 0; --st) {
                if (!yb_set[st]) {   // no stage combines R_st: its adjoint weight is zero
                    HC_HIP(hipMemsetAsync(s.yb[st], 0, len * sizeof(double), ctx->stream));
                    yb_set[st] = true;
                }
                // the adjoint residual of stage st is the adjoint of stage st - 1's output, delivered
                // straight into that stage's Runge-Kutta adjoint; stage 0's feeds the step's start
                if (st > 0) {
                    const RkDeliver prev = program(st - 1);
                    r = adjoint_at(ctx, s.seg[(size_t)t * ns + st], s.yb[st], nullptr, false, &prev);
                } else {
                    r = adjoint_at(ctx, s.seg[(size_t)t * ns], s.yb[0], nxt, true, nullptr);
                }
                if (r) return r;
            }
            std::swap(lam, nxt);
        }
    }
    if ((r = minf_total(ctx))) return r;
    if (dev) {
        if ((r = soa_to_aos_device(ctx, qbar_initial, lam, N, nv))) return r;
        HC_HIP(hipMemcpyAsync(bref_bar, s.brefbar_tot, nb * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        HC_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        if ((r = download_soa(ctx, qbar_initial, lam, N, nv))) return r;
        HC_HIP(hipMemcpy(bref_bar, s.brefbar_tot, nb * sizeof(double), hipMemcpyDeviceToHost));
    }
    HC_HIP(hipMemcpy(minf_bar, s.scal, sizeof(double), hipMemcpyDeviceToHost));
    return check_status(ctx);
}

int hc_adjoint_backward(hc_context* ctx, const double* qbar_final, double* qbar_initial, double* bref_bar,
                        double* minf_bar)
{
    return adjoint_backward(ctx, qbar_final, qbar_initial, bref_bar, minf_bar, false);
}

int hc_adjoint_backward_device(hc_context* ctx, const double* qbar_final_dev, double* qbar_initial_dev,
                               double* bref_bar_dev, double* minf_bar)
{
    return adjoint_backward(ctx, qbar_final_dev, qbar_initial_dev, bref_bar_dev, minf_bar, true);
}

int hc_tangent_steps(hc_context* ctx, int nsteps, const double* qdot0, const double* bref_dot, double minf_dot,
                     double* qdot_final)
{
    int r = adj_ready(ctx, true);
    if (r) return r;
    if (nsteps < 0 || !qdot0 || !qdot_final) return fail(ctx, HC_ERR_ARG, "invalid argument");
    if ((r = ensure(ctx))) return r;
    AdjState& s = *ctx->adj;
    const auto& c = ctx->cases[0];
    const int ns = s.ns, nv = s.nv;
    const long long N = s.N, len = (long long)nv * N;
    const double dt = ctx->cs.cfl;
    if ((r = grow(ctx, s.stage, ns, len)) || (r = grow(ctx, s.rd, ns, len)) || (r = grow(ctx, s.qdot, ns + 1, len)))
        return r;
    if ((r = upload_soa(ctx, s.qdot[0], qdot0, N, nv))) return r;
    if (bref_dot)
        HC_HIP(hipMemcpy(s.brefd_in, bref_dot, (size_t)c.brows * nv * sizeof(double), hipMemcpyHostToDevice));
    for (int t = 0; t < nsteps; ++t) {
        double* qn = ctx->q[ctx->cur];
        const double* in = qn;
        for (int st = 0; st < ns; ++st) {
            // tangent residual at (q_st, qdot_st)
            AdjArgs a = base_adj(ctx);
            a.q = in;
            a.qd = s.qdot[st];
            a.brefd = bref_dot ? s.brefd_in : nullptr;
            a.minfd = minf_dot;
            a.resd = s.rd[st];
            HC_HIP(run_tangent(ctx, a));
            RkTan tk{};
            tk.nk = st + 1;
            for (int k = 0; k <= st; ++k) {
                tk.rd[k] = s.rd[k];
                tk.coef[k] = ctx->cs.alpha[st][k];
            }
            hipLaunchKernelGGL(k_rk_tangent, dim3(blocks(len)), dim3(256), 0, ctx->stream, s.qdot[st + 1], s.qdot[0],
                               tk, ctx->cs.beta[st] * dt, ctx->cvol, N, nv);
            HC_HIP(hipGetLastError());
            // state: stage outputs through the fused kernel, the last one in place
            double* out = st == ns - 1 ? s.qtmp : s.stage[st + 1];
            if ((r = stage_into(ctx, st, in, qn, out))) return r;
            in = out;
        }
        if ((r = copy_d(ctx, qn, s.qtmp, len))) return r;
        if ((r = copy_d(ctx, s.qdot[0], s.qdot[ns], len))) return r;
    }
    return download_soa(ctx, qdot_final, s.qdot[0], N, nv);
}

}   // extern "C"
