// hc_capi.cpp -- the C-ABI of include/hcamd.h: context, its cases and device buffers, the
// per-step launch sequence of the stage kernels, tile-shape autotuning, HIP graph capture,
// norms.  A context holds one case (hc_set_mesh / hc_set_case) or a batch of cases
// (hc_batch_begin / hc_batch_add_case / hc_batch_end) that share the flux, limiter, basis
// and Runge-Kutta table; every stage launch covers the tiles of all its cases.
#include "hc_context.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

using namespace hc;

// n: slots (MAXSPREAD per case)
__global__ void k_reset_max(unsigned long long* p, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[(long long)k * MAXSTRIDE] = dbl_bits(DBL_MIN);
}

__global__ void k_clear_dt(double* p, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

__global__ void k_clear_u32(unsigned int* p) { *p = 0u; }

// Residual norms with wavefront reductions: each lane folds its cells (grid stride), each
// wavefront folds its 64 lanes with cross-lane shuffles (six exchanges, no LDS traffic), and one
// LDS exchange combines the four wavefronts of a work-group.  The order of every sum is fixed by
// the launch shape, so results are reproducible run to run.
__device__ inline double wave_sum_d(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_max_d(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// one partial per work-group: sums of (R/vol)^2 and max |R/vol| per variable, [8] per partial
__global__ void __launch_bounds__(256) k_norm_partials(const double* __restrict__ res, const double* __restrict__ vol,
                                                       long long N, int nv, double* __restrict__ part)
{
    __shared__ double w8[4][8];
    const int tid = threadIdx.x;
    double s[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
    for (long long c = (long long)blockIdx.x * 256 + tid; c < N; c += (long long)gridDim.x * 256) {
        const double vc = vol[c];
        for (int v = 0; v < nv; ++v) {
            const double r = res[c * (nv + 1) + v] / vc;
            s[v] = s[v] + r * r;
            m[v] = fmax(m[v], fabs(r));
        }
    }
    for (int v = 0; v < 4; ++v) {
        s[v] = wave_sum_d(s[v]);
        m[v] = wave_max_d(m[v]);
    }
    if ((tid & 63) == 0)
        for (int v = 0; v < 4; ++v) {
            w8[tid >> 6][v] = s[v];
            w8[tid >> 6][4 + v] = m[v];
        }
    __syncthreads();
    if (tid < 8) {
        double a = w8[0][tid];
        for (int w = 1; w < 4; ++w) a = tid < 4 ? a + w8[w][tid] : fmax(a, w8[w][tid]);
        part[blockIdx.x * 8 + tid] = a;
    }
}

// the case's norms from nparts partials: {L2[4] = sqrt(sum / N), Linf[4]}
__global__ void __launch_bounds__(256) k_norm_final(const double* __restrict__ part, int nparts, long long N,
                                                    double* __restrict__ out)
{
    __shared__ double w8[4][8];
    const int tid = threadIdx.x;
    double s[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
    for (int p = tid; p < nparts; p += 256)
        for (int v = 0; v < 4; ++v) {
            s[v] = s[v] + part[p * 8 + v];
            m[v] = fmax(m[v], part[p * 8 + 4 + v]);
        }
    for (int v = 0; v < 4; ++v) {
        s[v] = wave_sum_d(s[v]);
        m[v] = wave_max_d(m[v]);
    }
    if ((tid & 63) == 0)
        for (int v = 0; v < 4; ++v) {
            w8[tid >> 6][v] = s[v];
            w8[tid >> 6][4 + v] = m[v];
        }
    __syncthreads();
    if (tid < 8) {
        double a = w8[0][tid];
        for (int w = 1; w < 4; ++w) a = tid < 4 ? a + w8[w][tid] : fmax(a, w8[w][tid]);
        out[tid] = tid < 4 ? sqrt(a / (double)N) : a;
    }
}

constexpr int NORM_BLOCKS = 1024;

// STREAM triad on 16-byte elements: each lane issues the loads of four elements of b and c before
// any arithmetic (eight 16-byte loads in flight per lane), one pass over the arrays
constexpr int TRIAD_UNROLL = 4;
__global__ void __launch_bounds__(256) k_triad(double2* __restrict__ a, const double2* __restrict__ b,
                                               const double2* __restrict__ c, double s, long long n2)
{
    const long long base = (long long)blockIdx.x * 256 * TRIAD_UNROLL + threadIdx.x;
    double2 x[TRIAD_UNROLL], y[TRIAD_UNROLL];
#pragma unroll
    for (int u = 0; u < TRIAD_UNROLL; ++u) {
        const long long i = base + (long long)u * 256;
        if (i < n2) {
            x[u] = b[i];
            y[u] = c[i];
        }
    }
#pragma unroll
    for (int u = 0; u < TRIAD_UNROLL; ++u) {
        const long long i = base + (long long)u * 256;
        if (i < n2) a[i] = double2{x[u].x + s * y[u].x, x[u].y + s * y[u].y};
    }
}

// a = s b on 8-byte elements, one per lane: the stage kernels' access width, run once per triad
// call so that a counter pass (FETCH_SIZE / WRITE_SIZE) sees a kernel of known traffic at both widths
__global__ void __launch_bounds__(256) k_scale8(double* __restrict__ a, const double* __restrict__ b, double s,
                                                long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = s * b[i];
}

__global__ void __launch_bounds__(256) k_fill(double* a, double v, long long n)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) a[i] = v;
}

// [N][nv] (cell-major, as the C-ABI's host arrays) <-> [nv][N] planes, both in device memory
__global__ void __launch_bounds__(256) k_aos_to_soa(const double* __restrict__ aos, long long N, int nv,
                                                    double* __restrict__ soa)
{
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < N; c += (long long)gridDim.x * 256)
        for (int v = 0; v < nv; ++v) soa[v * N + c] = aos[c * nv + v];
}

__global__ void __launch_bounds__(256) k_soa_to_aos(const double* __restrict__ soa, long long N, int nv,
                                                    double* __restrict__ aos)
{
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < N; c += (long long)gridDim.x * 256)
        for (int v = 0; v < nv; ++v) aos[c * nv + v] = soa[v * N + c];
}

template <class T> void free_dev(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}   // namespace


namespace hc_detail {

// whether p is device (or managed) memory of this HIP runtime; a failed query's error is cleared
// so that it is not reported by a later hipGetLastError
bool device_pointer(const void* p)
{
    if (!p) return false;
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged || at.isManaged;
}

int aos_to_soa_device(hc_context* ctx, double* soa, const double* aos_dev, long long N, int nv)
{
    const long long nb = std::min<long long>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)std::max<long long>(nb, 1)), dim3(256), 0, ctx->stream, aos_dev, N,
                       nv, soa);
    HC_HIP(hipGetLastError());
    return HC_OK;
}

int soa_to_aos_device(hc_context* ctx, double* aos_dev, const double* soa, long long N, int nv)
{
    const long long nb = std::min<long long>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)std::max<long long>(nb, 1)), dim3(256), 0, ctx->stream, soa, N,
                       nv, aos_dev);
    HC_HIP(hipGetLastError());
    return HC_OK;
}

using namespace hc;

int fail(hc_context* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

int hipfail(hc_context* c, hipError_t e, const char* where)
{
    return fail(c, HC_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}


// Launches of the executable graph may still be in flight (hc_step returns without waiting), and
// HIP does not promise to defer the destruction of an executing graph: wait for the stream first.
void drop_graph(hc_context* ctx)
{
    if ((ctx->gexec || ctx->graph) && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->gexec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_cur = -1;
    ctx->graph_cfg = -1;
}

void release_buffers(hc_context* ctx)
{
    drop_graph(ctx);
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
    free_dev(ctx->maxslot);
    free_dev(ctx->dtslot);
    for (int k = 0; k < NTILECFG; ++k) {
        free_dev(ctx->rec64[k]);
        free_dev(ctx->rec32[k]);
        free_dev(ctx->tiles[k]);
        ctx->ntiles[k] = 0;
    }
    ctx->ngeo = 0;
    free_dev(ctx->chunks);
    ctx->nchunks = 0;
    free_dev(ctx->blk);
    ctx->blk_n = 0;
    free_dev(ctx->sm);
    ctx->sm_n = 0;
    ctx->smooth_mode = -1;
    ctx->smooth_alpha = 1.0;
    ctx->have_case = false;
    ctx->tuned = false;
    release_adjoint(ctx);
}

bool valid_ids(const hc_case_desc& d)
{
    if (d.law < 0 || d.law > 1) return false;
    if (d.basis < 0 || d.basis > 1) return false;
    if (d.law == HC_LAW_SCALAR && (d.basis != HC_BASIS_CONSERVED || d.flux > HC_FLUX_RUSANOV)) return false;
    if (d.flux < 0 || d.flux > 6) return false;
    if (d.law == HC_LAW_SCALAR && d.flux > HC_FLUX_RUSANOV) return false;
    if (d.limiter < 0 || d.limiter > 9) return false;
    if (d.vscale < 0 || d.vscale > 2 || d.pscale < 0 || d.pscale > 2) return false;
    for (int b = 0; b < 4; ++b) {
        if (d.bc[b] < 0 || d.bc[b] > 4) return false;
        if (d.law == HC_LAW_SCALAR && (d.bc[b] == HC_BC_ENTROPY || d.bc[b] == HC_BC_WALL)) return false;
    }
    if ((d.bc[0] == HC_BC_PERIODIC) != (d.bc[1] == HC_BC_PERIODIC)) return false;
    if ((d.bc[2] == HC_BC_PERIODIC) != (d.bc[3] == HC_BC_PERIODIC)) return false;
    if (d.nstages < 1 || d.nstages > 6) return false;
    if (d.dt_mode < 0 || d.dt_mode > 2) return false;
    if (!(d.cfl > 0.0)) return false;
    return true;
}

// the cases of one context share one kernel instantiation and one Runge-Kutta schedule
bool same_schedule(const hc_case_desc& a, const hc_case_desc& b)
{
    if (a.law != b.law || a.basis != b.basis || a.flux != b.flux || a.limiter != b.limiter) return false;
    if (a.vscale != b.vscale || a.pscale != b.pscale || a.nstages != b.nstages || a.dt_mode != b.dt_mode) return false;
    if (a.cfl != b.cfl) return false;
    for (int r = 0; r < 6; ++r) {
        if (a.beta[r] != b.beta[r]) return false;
        for (int k = 0; k < 6; ++k)
            if (a.alpha[r][k] != b.alpha[r][k]) return false;
    }
    return true;
}

int add_case(hc_context* ctx, int ni, int nj, const double* nodes, const hc_case_desc& desc)
{
    if (ni < 3 || nj < 3) return fail(ctx, HC_ERR_ARG, "mesh needs at least 3x3 cells");
    if ((long long)ni * nj > (1LL << 31) - 1) return fail(ctx, HC_ERR_ARG, "mesh too large");
    if (!valid_ids(desc)) return fail(ctx, HC_ERR_ARG, "invalid case descriptor");
    if (!ctx->cases.empty() && !same_schedule(ctx->cases[0].cs, desc))
        return fail(ctx, HC_ERR_ARG, "the cases of a batch must share law, basis, flux, limiter, scalings, "
                                     "Runge-Kutta table, time-step mode and CFL");
    CaseHost c;
    c.ni = ni;
    c.nj = nj;
    c.cs = desc;
    GeomInput gin{ni, nj, nodes, {desc.bc[0], desc.bc[1], desc.bc[2], desc.bc[3]},
                  {desc.period_i[0], desc.period_i[1]}, {desc.period_j[0], desc.period_j[1]}};
    if (build_geometry(gin, c.g) != 0) return fail(ctx, HC_ERR_ARG, "degenerate mesh cell");
    c.N = (long long)ni * nj;
    c.nfi = (long long)(ni + 1) * nj;
    c.nfj = (long long)ni * (nj + 1);
    c.brows = 2 * nj + 2 * ni;
    ctx->cases.push_back(std::move(c));
    return HC_OK;
}

template <class RF> void fill_rec(const hc_context* ctx, const CaseHost& c, int k, CaseRec<RF>& r)
{
    r = CaseRec<RF>{};
    r.m.ni = c.ni;
    r.m.nj = c.nj;
    r.m.per_i = c.cs.bc[0] == HC_BC_PERIODIC;
    r.m.per_j = c.cs.bc[2] == HC_BC_PERIODIC;
    r.m.nfi = c.nfi;
    r.m.nfj = c.nfj;
    r.m.cvol = ctx->cvol + c.geo_cell_off;
    r.m.ccx = ctx->ccx + c.geo_cell_off;
    r.m.ccy = ctx->ccy + c.geo_cell_off;
    if constexpr (sizeof(RF) == 8) {
        r.m.fi = ctx->fi + NSLOT * c.geo_fi_off;
        r.m.fj = ctx->fj + NSLOT * c.geo_fj_off;
    } else {
        r.m.fi = ctx->fi32 + NSLOT * c.geo_fi_off;
        r.m.fj = ctx->fj32 + NSLOT * c.geo_fj_off;
    }
    r.m.fi64 = ctx->fi + NSLOT * c.geo_fi_off;
    r.m.fj64 = ctx->fj + NSLOT * c.geo_fj_off;
    for (int d = 0; d < 2; ++d) {
        r.m.period_i[d] = c.cs.period_i[d];
        r.m.period_j[d] = c.cs.period_j[d];
    }
    const hc_species& s = c.cs.sp;
    r.sp = Species<double>{s.gamma, s.minf, s.lref, s.nu, s.pr, s.dt, s.R, s.gamma1};
    r.spf = Species<RF>{RF(s.gamma), RF(s.minf), RF(s.lref), RF(s.nu), RF(s.pr), RF(s.dt), RF(s.R), RF(s.gamma1)};
    for (int b = 0; b < 4; ++b) r.bc[b] = c.cs.bc[b];
    r.N = c.N;
    r.cell_off = c.cell_off;
    r.bref_off = c.bref_off;
    r.tiles_j = c.tiles_j[k];
}

// bitwise equality of two cases' device geometry (cells and face slots)
bool same_geometry(const GeomOutput& a, const GeomOutput& b)
{
    auto eq = [](const std::vector<double>& x, const std::vector<double>& y) {
        return x.size() == y.size() && (x.empty() || std::memcmp(x.data(), y.data(), x.size() * sizeof(double)) == 0);
    };
    return eq(a.cvol, b.cvol) && eq(a.ccx, b.ccx) && eq(a.ccy, b.ccy) && eq(a.fi, b.fi) && eq(a.fj, b.fj);
}

// Case records and tile tables of every tile shape.  The table of a shape holds the tiles of one
// step first: case by case, or with interleaving, every group of cases sharing a geometry block
// tile by tile (tile lt of each case of the group, then tile lt + 1 ...), so that the work-groups
// resident on an XCD together read the same face slots.  Each case's tiles follow in case order
// (the part a single-case launch covers, from tile0).
int build_tiles(hc_context* ctx)
{
    const TileShape* shapes = tile_shapes(ctx->precision == HC_PRECISION_MIXED, ctx->cs.law, ctx->cs.flux);
    const int nc = (int)ctx->cases.size();
    for (int k = 0; k < NTILECFG; ++k) {
        long long nt = 0;
        for (auto& c : ctx->cases) {
            const int ti = (c.ni + shapes[k].ti - 1) / shapes[k].ti;
            const int tj = (c.nj + shapes[k].tj - 1) / shapes[k].tj;
            c.tiles_j[k] = tj;
            c.ntiles[k] = ti * tj;
            nt += (long long)ti * tj;
        }
        if (2 * nt > (1LL << 31) - 1) return fail(ctx, HC_ERR_ARG, "too many tiles");
        std::vector<TileRef> tab;
        tab.reserve((size_t)(2 * nt));
        if (ctx->interleave) {
            std::vector<std::vector<int>> members(nc);   // cases of each group, keyed by its first case
            for (int ci = 0; ci < nc; ++ci) members[ctx->cases[ci].geo].push_back(ci);
            for (int g = 0; g < nc; ++g)
                for (int lt = 0; lt < ctx->cases[g].ntiles[k] && !members[g].empty(); ++lt)
                    for (int ci : members[g]) tab.push_back(TileRef{ci, lt});
        } else {
            for (int ci = 0; ci < nc; ++ci)
                for (int lt = 0; lt < ctx->cases[ci].ntiles[k]; ++lt) tab.push_back(TileRef{ci, lt});
        }
        for (int ci = 0; ci < nc; ++ci) {
            ctx->cases[ci].tile0[k] = (int)tab.size();
            for (int lt = 0; lt < ctx->cases[ci].ntiles[k]; ++lt) tab.push_back(TileRef{ci, lt});
        }
        std::vector<CaseRec<double>> r64(nc);
        std::vector<CaseRec<float>> r32(nc);
        for (int ci = 0; ci < nc; ++ci) {
            fill_rec(ctx, ctx->cases[ci], k, r64[ci]);
            fill_rec(ctx, ctx->cases[ci], k, r32[ci]);
        }
        free_dev(ctx->rec64[k]);
        free_dev(ctx->rec32[k]);
        free_dev(ctx->tiles[k]);
        HC_HIP(hipMalloc(&ctx->rec64[k], nc * sizeof(CaseRec<double>)));
        HC_HIP(hipMalloc(&ctx->rec32[k], nc * sizeof(CaseRec<float>)));
        HC_HIP(hipMalloc(&ctx->tiles[k], tab.size() * sizeof(TileRef)));
        HC_HIP(hipMemcpy(ctx->rec64[k], r64.data(), nc * sizeof(CaseRec<double>), hipMemcpyHostToDevice));
        HC_HIP(hipMemcpy(ctx->rec32[k], r32.data(), nc * sizeof(CaseRec<float>), hipMemcpyHostToDevice));
        HC_HIP(hipMemcpy(ctx->tiles[k], tab.data(), tab.size() * sizeof(TileRef), hipMemcpyHostToDevice));
        ctx->ntiles[k] = (int)nt;
    }
    return HC_OK;
}

// device layout of all cases: packed arrays, case records, tile tables, update chunks
int build_layout(hc_context* ctx)
{
    release_buffers(ctx);
    if (ctx->cases.empty()) return fail(ctx, HC_ERR_STATE, "no case");
    ctx->cs = ctx->cases[0].cs;
    ctx->nv = ctx->cs.law == HC_LAW_EULER ? 4 : 3;
    ctx->nslots = make_rk_plans(ctx->cs.nstages, ctx->cs.alpha, ctx->cs.beta, ctx->plans);
    const int nv = ctx->nv;
    const int nc = (int)ctx->cases.size();
    long long N = 0, nfi = 0, nfj = 0, br = 0;
    long long gN = 0, gfi = 0, gfj = 0;   // packed geometry: one block per group of equal geometries
    ctx->ngeo = 0;
    for (int ci = 0; ci < nc; ++ci) {
        CaseHost& c = ctx->cases[ci];
        c.cell_off = N;
        c.bref_off = br;
        N += c.N;
        nfi += c.nfi;
        nfj += c.nfj;
        br += c.brows;
        c.geo = ci;
        for (int g = 0; g < ci; ++g) {
            const CaseHost& o = ctx->cases[g];
            if (o.geo == g && o.ni == c.ni && o.nj == c.nj && same_geometry(o.g, c.g)) {
                c.geo = g;
                break;
            }
        }
        if (c.geo == ci) {
            c.geo_cell_off = gN;
            c.geo_fi_off = gfi;
            c.geo_fj_off = gfj;
            gN += c.N;
            gfi += c.nfi;
            gfj += c.nfj;
            ctx->ngeo++;
        } else {
            const CaseHost& o = ctx->cases[c.geo];
            c.geo_cell_off = o.geo_cell_off;
            c.geo_fi_off = o.geo_fi_off;
            c.geo_fj_off = o.geo_fj_off;
        }
    }
    ctx->Ntot = N;
    ctx->nfi_tot = nfi;
    ctx->nfj_tot = nfj;
    ctx->brows_tot = br;
    // geometry, one block per group
    HC_HIP(hipMalloc(&ctx->cvol, gN * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->ccx, gN * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->ccy, gN * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->fi, NSLOT * gfi * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->fj, NSLOT * gfj * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->fi32, NSLOT * gfi * sizeof(float)));
    HC_HIP(hipMalloc(&ctx->fj32, NSLOT * gfj * sizeof(float)));
    for (int ci = 0; ci < nc; ++ci) {
        CaseHost& c = ctx->cases[ci];
        if (c.geo == ci) {
            HC_HIP(hipMemcpy(ctx->cvol + c.geo_cell_off, c.g.cvol.data(), c.N * sizeof(double), hipMemcpyHostToDevice));
            HC_HIP(hipMemcpy(ctx->ccx + c.geo_cell_off, c.g.ccx.data(), c.N * sizeof(double), hipMemcpyHostToDevice));
            HC_HIP(hipMemcpy(ctx->ccy + c.geo_cell_off, c.g.ccy.data(), c.N * sizeof(double), hipMemcpyHostToDevice));
            HC_HIP(hipMemcpy(ctx->fi + NSLOT * c.geo_fi_off, c.g.fi.data(), c.g.fi.size() * sizeof(double),
                             hipMemcpyHostToDevice));
            HC_HIP(hipMemcpy(ctx->fj + NSLOT * c.geo_fj_off, c.g.fj.data(), c.g.fj.size() * sizeof(double),
                             hipMemcpyHostToDevice));
            std::vector<float> f32;
            fp32_slots(c.g.fi, c.nfi, f32);
            HC_HIP(hipMemcpy(ctx->fi32 + NSLOT * c.geo_fi_off, f32.data(), f32.size() * sizeof(float),
                             hipMemcpyHostToDevice));
            fp32_slots(c.g.fj, c.nfj, f32);
            HC_HIP(hipMemcpy(ctx->fj32 + NSLOT * c.geo_fj_off, f32.data(), f32.size() * sizeof(float),
                             hipMemcpyHostToDevice));
        }
        const GeomInput gin{c.ni, c.nj, nullptr, {c.cs.bc[0], c.cs.bc[1], c.cs.bc[2], c.cs.bc[3]}, {0.0, 0.0}, {0.0, 0.0}};
        c.uniform = uniform_geometry(gin, c.g);
        c.X = std::move(c.g.X);
    }
    for (auto& c : ctx->cases) c.g = GeomOutput{};   // the device holds it now
    // state, accumulators, per-case slots
    for (auto& b : ctx->q) {
        HC_HIP(hipMalloc(&b, (size_t)nv * N * sizeof(double)));
        HC_HIP(hipMemset(b, 0, (size_t)nv * N * sizeof(double)));
    }
    ctx->cur = 0;
    if (ctx->nslots > 0) HC_HIP(hipMalloc(&ctx->acc, (size_t)ctx->nslots * nv * N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->tmp, (size_t)nv * N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->lam0, N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->res, (size_t)(nv + 1) * N * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->bref, (size_t)br * nv * sizeof(double)));
    HC_HIP(hipMemset(ctx->bref, 0, (size_t)br * nv * sizeof(double)));
    HC_HIP(hipMalloc(&ctx->maxslot, (size_t)nc * MAXSPREAD * MAXSTRIDE * sizeof(unsigned long long)));
    HC_HIP(hipMalloc(&ctx->dtslot, 2 * nc * sizeof(double)));
    {
        std::vector<unsigned long long> m((size_t)nc * MAXSPREAD * MAXSTRIDE, dbl_bits(DBL_MIN));
        std::vector<double> z(2 * nc, 0.0);
        HC_HIP(hipMemcpy(ctx->maxslot, m.data(), m.size() * sizeof(m[0]), hipMemcpyHostToDevice));
        HC_HIP(hipMemcpy(ctx->dtslot, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // case records and tile tables per tile shape
    {
        const int r = build_tiles(ctx);
        if (r) return r;
    }
    // 256-cell chunks of the deferred global-dt update
    std::vector<UpdChunk> ch;
    for (int ci = 0; ci < nc; ++ci)
        for (long long s = 0; s < ctx->cases[ci].N; s += 256) ch.push_back(UpdChunk{ci, (int)s});
    HC_HIP(hipMalloc(&ctx->chunks, ch.size() * sizeof(UpdChunk)));
    HC_HIP(hipMemcpy(ctx->chunks, ch.data(), ch.size() * sizeof(UpdChunk), hipMemcpyHostToDevice));
    ctx->nchunks = (int)ch.size();
    ctx->uni_ok = true;
    for (const auto& c : ctx->cases) ctx->uni_ok = ctx->uni_ok && c.uniform;
    ctx->have_case = true;
    return HC_OK;
}

template <class RF> LaunchArgs<RF> launch_args(hc_context* ctx)
{
    LaunchArgs<RF> L{};
    if constexpr (sizeof(RF) == 8) L.cases = ctx->rec64[ctx->tile_cfg];
    else L.cases = ctx->rec32[ctx->tile_cfg];
    L.tiles = ctx->tiles[ctx->tile_cfg];
    L.tile_base = 0;
    L.ntiles = ctx->ntiles[ctx->tile_cfg];
    L.nslots = ctx->nslots;
    L.vscale = ctx->cs.vscale;
    L.pscale = ctx->cs.pscale;
    L.limiter = ctx->cs.limiter;
    L.dt_mode = ctx->cs.dt_mode;
    L.cfl = ctx->cs.cfl;
    L.acc = ctx->acc;
    L.tmp = ctx->tmp;
    L.lam0 = ctx->lam0;
    L.bref = ctx->bref;
    L.res_out = ctx->res;
    L.maxslot = ctx->maxslot;
    L.dtslot = ctx->dtslot;
    L.status = ctx->status;
    return L;
}

// the launch arguments differ between the precisions only in the case records
template <class RF> LaunchArgs<RF> with_cases(const LaunchArgs<double>& L, const CaseRec<RF>* recs)
{
    LaunchArgs<RF> o;
    std::memcpy(&o, &L, sizeof(L));
    static_assert(sizeof(LaunchArgs<RF>) == sizeof(LaunchArgs<double>), "layouts differ only in pointer types");
    o.cases = recs;
    return o;
}

hipError_t launch_stage(hc_context* ctx, const LaunchArgs<double>& L)
{
    const int key = ctx->cs.law * 100 + ctx->cs.basis * 10 + ctx->cs.flux;
    const int cfg = ctx->tile_cfg;
    const int uni = ctx->uni_ok && ctx->uni_enable ? 1 : 0;
    if (ctx->precision == HC_PRECISION_MIXED) {
        const LaunchArgs<float> L32 = with_cases<float>(L, ctx->rec32[cfg]);
#define HC_CASE_STAGE32(Lw, B, F) \
    case Lw * 100 + B * 10 + F: return hc_stage_L##Lw##_B##B##_F##F##_P32(L32, cfg, uni, ctx->stream);
        switch (key) {
            HC_FOR_EACH_KERNEL(HC_CASE_STAGE32)
        default: return hipErrorInvalidValue;
        }
#undef HC_CASE_STAGE32
    }
#define HC_CASE_STAGE64(Lw, B, F) \
    case Lw * 100 + B * 10 + F: return hc_stage_L##Lw##_B##B##_F##F##_P64(L, cfg, uni, ctx->stream);
    switch (key) {
        HC_FOR_EACH_KERNEL(HC_CASE_STAGE64)
    default: return hipErrorInvalidValue;
    }
#undef HC_CASE_STAGE64
}

hipError_t launch_update(hc_context* ctx, const LaunchArgs<double>& L)
{
    const int key = ctx->cs.law * 10 + ctx->cs.basis;
    if (ctx->precision == HC_PRECISION_MIXED) {
        const LaunchArgs<float> L32 = with_cases<float>(L, ctx->rec32[ctx->tile_cfg]);
#define HC_CASE_UPD32(Lw, B) \
    case Lw * 10 + B: return hc_upd_L##Lw##_B##B##_P32(L32, ctx->chunks, ctx->nchunks, ctx->stream);
        switch (key) {
            HC_FOR_EACH_LAWBASIS(HC_CASE_UPD32)
        default: return hipErrorInvalidValue;
        }
#undef HC_CASE_UPD32
    }
#define HC_CASE_UPD64(Lw, B) \
    case Lw * 10 + B: return hc_upd_L##Lw##_B##B##_P64(L, ctx->chunks, ctx->nchunks, ctx->stream);
    switch (key) {
        HC_FOR_EACH_LAWBASIS(HC_CASE_UPD64)
    default: return hipErrorInvalidValue;
    }
#undef HC_CASE_UPD64
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

// the launches of one time step of every case (timestepping/rungeKutta.h:97-164)
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
    LaunchArgs<double> L = launch_args<double>(ctx);
    const double* in = q0;
    for (int s = 0; s < ns; ++s) {
        double* out;
        if (ns == 1) out = bufA;
        else if (s == ns - 1) out = q0;   // the last stage updates the step-start state in place
        else out = (s % 2 == 0) ? bufA : bufB;
        L.q_in = in;
        L.q0 = q0;
        L.q_out = out;
        L.stage = s;
        L.plan = ctx->plans[s];
        L.reset_max = (global && ns >= 2 && s == ns - 1) ? 1 : 0;
        L.mode = (global && s == 0) ? MODE_GLOBAL_STAGE0 : MODE_FUSED;
        L.keep_res = (ctx->keep_res && s == 0) ? 1 : 0;
        hipError_t e = timed(ctx, 0, [&] { return launch_stage(ctx, L); });
        if (e != hipSuccess) return hipfail(ctx, e, "stage kernel launch");
        if (global && s == 0) {
            e = timed(ctx, 1, [&] { return launch_update(ctx, L); });
            if (e != hipSuccess) return hipfail(ctx, e, "update kernel launch");
        }
        in = out;
    }
    if (global && ns == 1) {
        const int nc = (int)ctx->cases.size();
        const int nslot = nc * MAXSPREAD;
        hipLaunchKernelGGL(k_reset_max, dim3((nslot + 255) / 256), dim3(256), 0, ctx->stream, ctx->maxslot, nslot);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hipfail(ctx, e, "reset kernel launch");
    }
    if (ns == 1) ctx->cur = 1 - ctx->cur;
    return HC_OK;
}

int ready(hc_context* ctx)
{
    if (!ctx) return HC_ERR_ARG;
    if (ctx->building) return fail(ctx, HC_ERR_STATE, "batch not finished (hc_batch_end)");
    if (!ctx->have_case) return fail(ctx, HC_ERR_STATE, "mesh and case must be set first");
    return HC_OK;
}

int case_ready(hc_context* ctx, int k)
{
    int r = ready(ctx);
    if (r) return r;
    if (k < 0 || k >= (int)ctx->cases.size()) return fail(ctx, HC_ERR_ARG, "case index out of range");
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

int set_state(hc_context* ctx, int k, const double* q)
{
    int r = case_ready(ctx, k);
    if (r) return r;
    if (!q) return fail(ctx, HC_ERR_ARG, "null state");
    CaseHost& c = ctx->cases[k];
    const int nv = ctx->nv;
    std::vector<double> soa((size_t)nv * c.N);
    for (long long i = 0; i < c.N; ++i)
        for (int v = 0; v < nv; ++v) soa[(size_t)v * c.N + i] = q[i * nv + v];
    HC_HIP(hipStreamSynchronize(ctx->stream));
    HC_HIP(hipMemcpy(ctx->q[ctx->cur] + nv * c.cell_off, soa.data(), soa.size() * sizeof(double),
                     hipMemcpyHostToDevice));
    const double z[2] = {0.0, 0.0};
    HC_HIP(hipMemcpy(ctx->dtslot + 2 * k, z, sizeof(z), hipMemcpyHostToDevice));
    c.have_state = true;
    return HC_OK;
}

int get_state(hc_context* ctx, int k, double* q)
{
    int r = case_ready(ctx, k);
    if (r) return r;
    const CaseHost& c = ctx->cases[k];
    if (!q || !c.have_state) return fail(ctx, HC_ERR_STATE, "no state");
    const int nv = ctx->nv;
    std::vector<double> soa((size_t)nv * c.N);
    HC_HIP(hipStreamSynchronize(ctx->stream));
    HC_HIP(hipMemcpy(soa.data(), ctx->q[ctx->cur] + nv * c.cell_off, soa.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (long long i = 0; i < c.N; ++i)
        for (int v = 0; v < nv; ++v) q[i * nv + v] = soa[(size_t)v * c.N + i];
    return HC_OK;
}

int set_boundary_refs(hc_context* ctx, int k, const double* bref)
{
    int r = case_ready(ctx, k);
    if (r) return r;
    if (!bref) return fail(ctx, HC_ERR_ARG, "null boundary reference array");
    const CaseHost& c = ctx->cases[k];
    const size_t nb = (size_t)c.brows * ctx->nv;
    HC_HIP(hipMemcpyAsync(ctx->bref + ctx->nv * c.bref_off, bref, nb * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

// residual of case k into ctx->res (only its tiles are launched)
int residual_to_device(hc_context* ctx, int k)
{
    const CaseHost& c = ctx->cases[k];
    if (!c.have_state) return fail(ctx, HC_ERR_STATE, "no state");
    LaunchArgs<double> L = launch_args<double>(ctx);
    L.mode = MODE_RESIDUAL;
    L.q_in = ctx->q[ctx->cur];
    L.q0 = ctx->q[ctx->cur];
    L.q_out = ctx->q[ctx->cur];   // not written in residual mode
    L.plan = ctx->plans[0];
    L.tile_base = c.tile0[ctx->tile_cfg];
    L.ntiles = c.ntiles[ctx->tile_cfg];
    const hipError_t e = launch_stage(ctx, L);
    if (e != hipSuccess) return hipfail(ctx, e, "residual kernel launch");
    return HC_OK;
}

int residual(hc_context* ctx, int k, double* res)
{
    int r = case_ready(ctx, k);
    if (r) return r;
    if (!res) return fail(ctx, HC_ERR_ARG, "null residual array");
    r = residual_to_device(ctx, k);
    if (r) return r;
    const CaseHost& c = ctx->cases[k];
    HC_HIP(hipMemcpyAsync(res, ctx->res + (ctx->nv + 1) * c.cell_off, (size_t)(ctx->nv + 1) * c.N * sizeof(double),
                          hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return check_status(ctx);
}

// norms of the residual of case k held in ctx->res
int norms_of_res(hc_context* ctx, int k, double* l2, double* linf)
{
    const CaseHost& c = ctx->cases[k];
    hipLaunchKernelGGL(k_norm_partials, dim3(NORM_BLOCKS), dim3(256), 0, ctx->stream,
                       ctx->res + (ctx->nv + 1) * c.cell_off, ctx->cvol + c.geo_cell_off, c.N, ctx->nv, ctx->part);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, ctx->stream, ctx->part, NORM_BLOCKS, c.N, ctx->nout);
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

int norms(hc_context* ctx, int k, double* l2, double* linf)
{
    int r = case_ready(ctx, k);
    if (r) return r;
    if (!l2 || !linf) return fail(ctx, HC_ERR_ARG, "null output");
    r = residual_to_device(ctx, k);
    if (r) return r;
    return norms_of_res(ctx, k, l2, linf);
}

int get_time(hc_context* ctx, int k, double* t)
{
    int r = case_ready(ctx, k);
    if (r) return r;
    if (!t) return fail(ctx, HC_ERR_ARG, "null output");
    double d[2];
    HC_HIP(hipMemcpyAsync(d, ctx->dtslot + 2 * k, sizeof(d), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    *t = d[1];
    return HC_OK;
}

// steps per shape and repetitions of the tile-shape trial
constexpr int TUNE_STEPS = 1, TUNE_REPS = 3;

// time every tile shape on real steps (the shapes give bit-identical results) and keep the fastest
int autotune(hc_context* ctx)
{
    double best[NTILECFG];
    for (int k = 0; k < NTILECFG; ++k) best[k] = 1e300;
    hipEvent_t e0 = take_event(ctx), e1 = take_event(ctx);
    for (int rep = 0; rep < TUNE_REPS; ++rep)
        for (int k = 0; k < NTILECFG; ++k) {
            ctx->tile_cfg = k;
            (void)hipEventRecord(e0, ctx->stream);
            for (int s = 0; s < TUNE_STEPS; ++s) {
                const int r = enqueue_step(ctx);
                if (r) return r;
            }
            (void)hipEventRecord(e1, ctx->stream);
            HC_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            HC_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best[k]) best[k] = ms;
        }
    ctx->ev_pool.push_back(e0);
    ctx->ev_pool.push_back(e1);
    int arg = 0;
    for (int k = 0; k < NTILECFG; ++k) {
        ctx->tune_ms[k] = best[k] / TUNE_STEPS;
        if (best[k] < best[arg]) arg = k;
    }
    ctx->tile_cfg = arg;
    ctx->tuned = true;
    return HC_OK;
}

int all_states(hc_context* ctx)
{
    for (const auto& c : ctx->cases)
        if (!c.have_state) return fail(ctx, HC_ERR_STATE, "no state");
    return HC_OK;
}

template hc::LaunchArgs<double> launch_args<double>(hc_context* ctx);

}   // namespace hc_detail

using namespace hc_detail;

extern "C" {

const char* hc_version(void) { return "hyperboliconvergence_amd 0.2 (gfx950)"; }

int hc_create(hc_context** out, int device)
{
    if (!out) return HC_ERR_ARG;
    *out = nullptr;
    hc_context* ctx = new hc_context();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&ctx->status, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMalloc(&ctx->part, NORM_BLOCKS * 8 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->nout, 8 * sizeof(double));
    if (e == hipSuccess) {
        const unsigned int zs = 0;
        e = hipMemcpy(ctx->status, &zs, sizeof(zs), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess || noalloc_setup(device) != HC_OK) {
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
    ctx->cases.clear();
    ctx->building = false;
    ctx->ni = ni;
    ctx->nj = nj;
    ctx->nodes.assign(nodes, nodes + 2 * (size_t)(ni + 1) * (nj + 1));
    ctx->have_mesh = true;
    return HC_OK;
}

int hc_set_case(hc_context* ctx, const hc_case_desc* desc)
{
    if (!ctx || !desc) return HC_ERR_ARG;
    if (!ctx->have_mesh) return fail(ctx, HC_ERR_STATE, "hc_set_mesh must come first");
    if (!valid_ids(*desc)) return fail(ctx, HC_ERR_ARG, "invalid case descriptor");
    HC_HIP(hipSetDevice(ctx->device));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    // keep the state and the boundary references across a case change with the same variable count
    std::vector<double> keep, keep_bref;
    const int nv_new = desc->law == HC_LAW_EULER ? 4 : 3;
    if (ctx->have_case && ctx->cases.size() == 1 && ctx->nv == nv_new) {
        const CaseHost& c = ctx->cases[0];
        if (c.have_state) {
            keep.resize((size_t)nv_new * c.N);
            HC_HIP(hipMemcpy(keep.data(), ctx->q[ctx->cur], keep.size() * sizeof(double), hipMemcpyDeviceToHost));
        }
        keep_bref.resize((size_t)c.brows * nv_new);
        HC_HIP(hipMemcpy(keep_bref.data(), ctx->bref, keep_bref.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    release_buffers(ctx);
    ctx->cases.clear();
    int r = add_case(ctx, ctx->ni, ctx->nj, ctx->nodes.data(), *desc);
    if (r) return r;
    r = build_layout(ctx);
    if (r) return r;
    if (!keep_bref.empty())
        HC_HIP(hipMemcpy(ctx->bref, keep_bref.data(), keep_bref.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!keep.empty()) {
        HC_HIP(hipMemcpy(ctx->q[0], keep.data(), keep.size() * sizeof(double), hipMemcpyHostToDevice));
        ctx->cases[0].have_state = true;
    }
    return HC_OK;
}

int hc_batch_begin(hc_context* ctx)
{
    if (!ctx) return HC_ERR_ARG;
    HC_HIP(hipSetDevice(ctx->device));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    release_buffers(ctx);
    ctx->cases.clear();
    ctx->have_mesh = false;
    ctx->building = true;
    return HC_OK;
}

int hc_batch_add_case(hc_context* ctx, int ni, int nj, const double* nodes, const hc_case_desc* desc, int* case_id)
{
    if (!ctx || !nodes || !desc) return HC_ERR_ARG;
    if (!ctx->building) return fail(ctx, HC_ERR_STATE, "hc_batch_begin must come first");
    const int r = add_case(ctx, ni, nj, nodes, *desc);
    if (r) return r;
    if (case_id) *case_id = (int)ctx->cases.size() - 1;
    return HC_OK;
}

int hc_batch_end(hc_context* ctx)
{
    if (!ctx) return HC_ERR_ARG;
    if (!ctx->building) return fail(ctx, HC_ERR_STATE, "hc_batch_begin must come first");
    HC_HIP(hipSetDevice(ctx->device));
    ctx->building = false;
    return build_layout(ctx);
}

int hc_batch_size(hc_context* ctx, int* ncases)
{
    if (!ctx || !ncases) return HC_ERR_ARG;
    *ncases = (int)ctx->cases.size();
    return HC_OK;
}

int hc_set_precision(hc_context* ctx, int precision)
{
    if (!ctx || (precision != HC_PRECISION_FP64 && precision != HC_PRECISION_MIXED)) return HC_ERR_ARG;
    if (ctx->precision == precision) return HC_OK;
    if (ctx->building) return fail(ctx, HC_ERR_STATE, "batch not finished (hc_batch_end)");
    ctx->precision = precision;
    if (!ctx->have_case) return HC_OK;
    // the tile tables depend on the precision's tile shapes: rebuild them, keeping the cases' data
    HC_HIP(hipSetDevice(ctx->device));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    const int r = build_tiles(ctx);
    if (r) return r;
    ctx->tuned = false;
    ctx->tile_cfg = ctx->tile_mode >= 0 ? ctx->tile_mode : 0;
    return HC_OK;
}

int hc_set_boundary_refs(hc_context* ctx, const double* bref) { return set_boundary_refs(ctx, 0, bref); }

int hc_set_state_device(hc_context* ctx, const double* q_dev)
{
    int r = case_ready(ctx, 0);
    if (r) return r;
    if (!device_pointer(q_dev)) return fail(ctx, HC_ERR_ARG, "q_dev is not device memory of the library's HIP runtime");
    CaseHost& c = ctx->cases[0];
    HC_HIP(hipStreamSynchronize(ctx->stream));
    if ((r = aos_to_soa_device(ctx, ctx->q[ctx->cur] + ctx->nv * c.cell_off, q_dev, c.N, ctx->nv))) return r;
    const double z[2] = {0.0, 0.0};
    HC_HIP(hipMemcpyAsync(ctx->dtslot, z, sizeof(z), hipMemcpyHostToDevice, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    c.have_state = true;
    return HC_OK;
}

int hc_get_state_device(hc_context* ctx, double* q_dev)
{
    int r = case_ready(ctx, 0);
    if (r) return r;
    const CaseHost& c = ctx->cases[0];
    if (!c.have_state) return fail(ctx, HC_ERR_STATE, "no state");
    if (!device_pointer(q_dev)) return fail(ctx, HC_ERR_ARG, "q_dev is not device memory of the library's HIP runtime");
    if ((r = soa_to_aos_device(ctx, q_dev, ctx->q[ctx->cur] + ctx->nv * c.cell_off, c.N, ctx->nv))) return r;
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_set_boundary_refs_device(hc_context* ctx, const double* bref_dev)
{
    int r = case_ready(ctx, 0);
    if (r) return r;
    if (!device_pointer(bref_dev))
        return fail(ctx, HC_ERR_ARG, "bref_dev is not device memory of the library's HIP runtime");
    const CaseHost& c = ctx->cases[0];
    HC_HIP(hipMemcpyAsync(ctx->bref + ctx->nv * c.bref_off, bref_dev, (size_t)c.brows * ctx->nv * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}
int hc_set_state(hc_context* ctx, const double* q) { return set_state(ctx, 0, q); }
int hc_get_state(hc_context* ctx, double* q) { return get_state(ctx, 0, q); }
int hc_residual(hc_context* ctx, double* res) { return residual(ctx, 0, res); }
int hc_norms(hc_context* ctx, double* l2, double* linf) { return norms(ctx, 0, l2, linf); }
int hc_get_time(hc_context* ctx, double* t) { return get_time(ctx, 0, t); }

int hc_batch_set_boundary_refs(hc_context* ctx, int k, const double* bref) { return set_boundary_refs(ctx, k, bref); }
int hc_batch_set_state(hc_context* ctx, int k, const double* q) { return set_state(ctx, k, q); }
int hc_batch_get_state(hc_context* ctx, int k, double* q) { return get_state(ctx, k, q); }
int hc_batch_residual(hc_context* ctx, int k, double* res) { return residual(ctx, k, res); }
int hc_batch_norms(hc_context* ctx, int k, double* l2, double* linf) { return norms(ctx, k, l2, linf); }
int hc_batch_get_time(hc_context* ctx, int k, double* t) { return get_time(ctx, k, t); }

int hc_step_norms(hc_context* ctx, int nsteps, double* l2, double* linf)
{
    int r = ready(ctx);
    if (r) return r;
    if (nsteps < 1 || !l2 || !linf) return fail(ctx, HC_ERR_ARG, "hc_step_norms: nsteps >= 1 and outputs");
    if (ctx->cases.size() != 1) return fail(ctx, HC_ERR_STATE, "hc_step_norms drives a single-case context");
    if ((r = hc_step(ctx, nsteps - 1))) return r;
    if (ctx->smooth_mode >= 0) {
        // the smoothed step has no fused residual store: the same residual, evaluated once more
        if ((r = hc_norms(ctx, l2, linf))) return r;
        return hc_step(ctx, 1);
    }
    if ((r = all_states(ctx))) return r;
    ctx->keep_res = true;   // the last step, eager: its stage 0 stores the residual of its start state
    r = enqueue_step(ctx);
    ctx->keep_res = false;
    if (r) return r;
    return norms_of_res(ctx, 0, l2, linf);
}

int hc_step_steady(hc_context* ctx, long long max_steps, double residual_drop, int check_every,
                   long long* steps_done, double* drop)
{
    int r = ready(ctx);
    if (r) return r;
    if (max_steps < 0 || check_every < 1 || !(residual_drop > 0.0) || !steps_done || !drop)
        return fail(ctx, HC_ERR_ARG, "steady stepping: max_steps >= 0, check_every >= 1, residual_drop > 0 and outputs");
    if (ctx->cases.size() != 1) return fail(ctx, HC_ERR_STATE, "steady stepping drives a single-case context");
    *steps_done = 0;
    *drop = 1.0;
    double l2[8], linf[8];
    r = hc_norms(ctx, l2, linf);
    if (r) return r;
    const double r0 = l2[2];
    if (!(r0 > 0.0)) return HC_OK;
    while (*steps_done < max_steps) {
        const int n = (int)std::min<long long>(check_every, max_steps - *steps_done);
        r = hc_step(ctx, n);
        if (r) return r;
        *steps_done += n;
        r = hc_norms(ctx, l2, linf);
        if (r) return r;
        *drop = l2[2] / r0;
        if (*drop <= residual_drop) break;
    }
    return HC_OK;
}

int hc_step(hc_context* ctx, int nsteps)
{
    int r = ready(ctx);
    if (r) return r;
    if (nsteps < 0) return fail(ctx, HC_ERR_ARG, "negative step count");
    r = all_states(ctx);
    if (r) return r;
    int done = 0;
    if (ctx->smooth_mode >= 0) {
        // residual smoothing: the unfused step (no graphs, the tile shape in use)
        for (; done < nsteps; ++done) {
            r = enqueue_step_smoothed(ctx);
            if (r) return r;
        }
        return HC_OK;
    }
    hipEvent_t r0 = nullptr;
    if (ctx->profiling == 2 && nsteps > 0) {
        r0 = take_event(ctx);
        (void)hipEventRecord(r0, ctx->stream);
    }
    // tile shape: fixed, or timed on the first steps of the context
    if (ctx->tile_mode >= 0) {
        ctx->tile_cfg = ctx->tile_mode;
    } else if (!ctx->tuned && nsteps - done >= NTILECFG * TUNE_REPS * TUNE_STEPS && ctx->profiling != 1) {
        r = autotune(ctx);
        if (r) return r;
        done += NTILECFG * TUNE_REPS * TUNE_STEPS;
    }
    int k = ctx->graph_steps;
    // single-stage tables alternate two state buffers: capture an even number of steps
    if (ctx->cs.nstages == 1 && (k % 2)) k += 1;
    if (k > 0 && ctx->profiling != 1) {
        if (ctx->gexec && (ctx->graph_cur != ctx->cur || ctx->graph_cfg != ctx->tile_cfg)) drop_graph(ctx);
        while (nsteps - done >= k) {
            if (!ctx->gexec) {
                const int cur0 = ctx->cur;
                HC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                int rc = HC_OK;
                for (int s = 0; s < k && rc == HC_OK; ++s) rc = enqueue_step(ctx);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
                ctx->cur = cur0;
                if (rc != HC_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (ec != hipSuccess) return hipfail(ctx, ec, "hipStreamEndCapture");
                ctx->graph = g;
                HC_HIP(hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0));
                ctx->graph_cur = cur0;
                ctx->graph_cfg = ctx->tile_cfg;
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
        const double upd = ctx->cs.dt_mode == HC_DT_GLOBAL ? 1.0 : 0.0;
        ctx->ev_pending.push_back({r0, r1, double(nsteps) * ctx->cs.nstages, double(nsteps) * upd, 2});
    }
    if (ctx->profiling == 1) drain_events(ctx);
    return HC_OK;
}

int hc_set_tiles(hc_context* ctx, int config)
{
    if (!ctx) return HC_ERR_ARG;
    if (config < -1 || config >= NTILECFG) return fail(ctx, HC_ERR_ARG, "tile configuration out of range");
    ctx->tile_mode = config;
    ctx->tuned = false;
    ctx->tile_cfg = config >= 0 ? config : 0;
    return HC_OK;
}

int hc_get_tiles(hc_context* ctx, int* config, double* ms_per_step)
{
    if (!ctx || !config) return HC_ERR_ARG;
    *config = ctx->tile_cfg;
    if (ms_per_step)
        for (int k = 0; k < NTILECFG; ++k) ms_per_step[k] = ctx->tuned ? ctx->tune_ms[k] : 0.0;
    return HC_OK;
}

int hc_set_uniform_geometry(hc_context* ctx, int enable)
{
    if (!ctx) return HC_ERR_ARG;
    if (enable != 0 && enable != 1) return fail(ctx, HC_ERR_ARG, "enable must be 0 or 1");
    if (ctx->uni_enable != (enable != 0)) {
        drop_graph(ctx);       // captured launches name the other kernels
        ctx->tuned = false;    // and their timings differ
    }
    ctx->uni_enable = enable != 0;
    return HC_OK;
}

int hc_get_uniform_geometry(hc_context* ctx, int* active)
{
    if (!ctx || !active) return HC_ERR_ARG;
    *active = ctx->have_case && ctx->uni_ok && ctx->uni_enable ? 1 : 0;
    return HC_OK;
}

int hc_set_tile_order(hc_context* ctx, int interleave)
{
    if (!ctx) return HC_ERR_ARG;
    if (ctx->building) return fail(ctx, HC_ERR_STATE, "batch not finished (hc_batch_end)");
    if (ctx->interleave == (interleave != 0)) return HC_OK;
    ctx->interleave = interleave != 0;
    if (!ctx->have_case) return HC_OK;
    HC_HIP(hipSetDevice(ctx->device));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    const int r = build_tiles(ctx);
    if (r) return r;
    ctx->tuned = false;
    ctx->tile_cfg = ctx->tile_mode >= 0 ? ctx->tile_mode : 0;
    return HC_OK;
}

int hc_get_geometry_blocks(hc_context* ctx, int* nblocks)
{
    if (!ctx || !nblocks) return HC_ERR_ARG;
    *nblocks = ctx->have_case ? ctx->ngeo : 0;
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
    if (k != ctx->graph_steps) drop_graph(ctx);
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
    if (!ctx || n < 2 || reps <= 0 || !gbps) return HC_ERR_ARG;
    HC_HIP(hipSetDevice(ctx->device));
    // released on every return path
    struct Scratch {
        double *a = nullptr, *b = nullptr, *c = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Scratch()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipFree(a);
            (void)hipFree(b);
            (void)hipFree(c);
        }
    } w;
    HC_HIP(hipMalloc(&w.a, n * sizeof(double)));
    HC_HIP(hipMalloc(&w.b, n * sizeof(double)));
    HC_HIP(hipMalloc(&w.c, n * sizeof(double)));
    const unsigned fill_grid = 256 * 32;
    hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, ctx->stream, w.b, 1.0, n);
    hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, ctx->stream, w.c, 2.0, n);
    const long long n2 = n / 2;   // 16-byte elements (an odd last double is not part of the measure)
    const long long per_block = 256LL * TRIAD_UNROLL;
    const unsigned grid = (unsigned)((n2 + per_block - 1) / per_block);
    double2* a2 = reinterpret_cast<double2*>(w.a);
    const double2* b2 = reinterpret_cast<const double2*>(w.b);
    const double2* c2 = reinterpret_cast<const double2*>(w.c);
    hipLaunchKernelGGL(k_triad, dim3(grid), dim3(256), 0, ctx->stream, a2, b2, c2, 3.0, n2);   // warm-up
    HC_HIP(hipGetLastError());
    HC_HIP(hipEventCreate(&w.e0));
    HC_HIP(hipEventCreate(&w.e1));
    HC_HIP(hipEventRecord(w.e0, ctx->stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_triad, dim3(grid), dim3(256), 0, ctx->stream, a2, b2, c2, 3.0, n2);
    HC_HIP(hipGetLastError());
    HC_HIP(hipEventRecord(w.e1, ctx->stream));
    HC_HIP(hipEventSynchronize(w.e1));
    float ms = 0.f;
    HC_HIP(hipEventElapsedTime(&ms, w.e0, w.e1));
    *gbps = 48.0 * (double)n2 * reps / (ms * 1e-3) / 1e9;
    hipLaunchKernelGGL(k_scale8, dim3((unsigned)((2 * n2 + 255) / 256)), dim3(256), 0, ctx->stream, w.a, w.b, 3.0,
                       2 * n2);
    HC_HIP(hipGetLastError());
    HC_HIP(hipStreamSynchronize(ctx->stream));
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
