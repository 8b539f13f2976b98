// hc_spatial.hip -- the reference's spatial building blocks as device kernels on a single-case
// context: Q metric, ghost-value update, spectral radius and the forward-Euler updates.  They
// serve callers that run the time loop themselves on host arrays (the C++ drop-in's xmetrics /
// qmetrics / boundaryUpdate / spectralRadius / eulerForwardUpdateGlobal / Local); hc_step runs
// the same arithmetic fused into the stage kernel.  Host arrays are staged through the context's
// device buffers between steps (those buffers hold no step state then).  Compiled with
// -ffp-contract=off: each result is bit-identical to the fused kernel's and to the oracle's.
#include "hc_context.hpp"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <mutex>
#include <vector>

namespace {

using namespace hc;

// Q metric of every cell (spatial/lsqMetrics.h:181-213; lsq/qmetric.ipp:48-70): per cell the
// interior i-edges, the interior j-edges, then the periodic wraps, dq = q_R - q_L times the
// centre difference of the face slots (the stage kernel's qmetric phase, in the same order)
template <int LAW>
__global__ void __launch_bounds__(256) k_qmetrics(const MeshView<double> m, const double* __restrict__ q, long long N,
                                                  double* __restrict__ Q)
{
    constexpr int NV = LawTraits<LAW>::nv;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const int ni = m.ni, nj = m.nj;
    const int i = (int)(c / nj), j = (int)(c % nj);
    double acc[NV][2];
    for (int v = 0; v < NV; ++v) acc[v][0] = acc[v][1] = 0.0;
    auto edge = [&](long long cl, long long cr, const double* base, long long nf, long long fidx) {
        const double dx0 = base[FS_DX0 * nf + fidx], dx1 = base[FS_DX1 * nf + fidx];
        for (int v = 0; v < NV; ++v) {
            const double dq = q[cr * NV + v] - q[cl * NV + v];
            acc[v][0] = acc[v][0] + dq * dx0;
            acc[v][1] = acc[v][1] + dq * dx1;
        }
    };
    const long long up = (long long)(i - 1) * nj + j, dn = (long long)(i + 1) * nj + j;
    if (i > 0) edge(up, c, m.fi, m.nfi, (long long)i * nj + j);
    if (i < ni - 1) edge(c, dn, m.fi, m.nfi, (long long)(i + 1) * nj + j);
    if (j > 0) edge(c - 1, c, m.fj, m.nfj, (long long)i * (nj + 1) + j);
    if (j < nj - 1) edge(c, c + 1, m.fj, m.nfj, (long long)i * (nj + 1) + j + 1);
    if (m.per_i) {
        if (i == 0) edge((long long)(ni - 1) * nj + j, c, m.fi, m.nfi, (long long)j);
        if (i == ni - 1) edge(c, (long long)j, m.fi, m.nfi, (long long)j);
    }
    if (m.per_j) {
        if (j == 0) edge((long long)i * nj + nj - 1, c, m.fj, m.nfj, (long long)i * (nj + 1));
        if (j == nj - 1) edge(c, (long long)i * nj, m.fj, m.nfj, (long long)i * (nj + 1));
    }
    for (int v = 0; v < NV; ++v) {
        Q[(c * NV + v) * 2 + 0] = acc[v][0];
        Q[(c * NV + v) * 2 + 1] = acc[v][1];
    }
}

// ghost values of the boundary cells (spatial/boundary/boundaryUpdate.h:28-48, 137-266): row r of
// [2 nj + 2 ni] (boundaries 0, 1 along j, then 2, 3 along i); the face normal points into the
// domain; Riemann / entropy far field from the point functions, Fixed and the inviscid wall
// return the reference value (boundaryCondition.h:63-80, inviscidWall.ipp:15-28); periodic rows
// are not written
template <int LAW, int BASIS>
__global__ void __launch_bounds__(256) k_boundary_update(const MeshView<double> m, Species<double> sp, int4 bc,
                                                         const double* __restrict__ q, const double* __restrict__ bref,
                                                         double* __restrict__ ghost, unsigned int* status)
{
    constexpr int NV = LawTraits<LAW>::nv;
    const int ni = m.ni, nj = m.nj;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= 2 * nj + 2 * ni) return;
    int b, t;
    if (r < 2 * nj) {
        b = r < nj ? 0 : 1;
        t = r - b * nj;
    } else {
        b = r - 2 * nj < ni ? 2 : 3;
        t = r - 2 * nj - (b - 2) * ni;
    }
    const int type = b == 0 ? bc.x : b == 1 ? bc.y : b == 2 ? bc.z : bc.w;
    if (type == BC_PERIODIC) return;
    double* out = ghost + (long long)r * NV;
    const double* qref = bref + (long long)r * NV;
    if (type == BC_WALL) {
        for (int v = 0; v < NV; ++v) out[v] = qref[v];
        return;
    }
    const bool isi = b < 2;
    const int l = isi ? (b == 0 ? 0 : ni) : (b == 2 ? 0 : nj);            // node line of the face
    const long long fidx = isi ? (long long)l * nj + t : (long long)t * (nj + 1) + l;
    const double* base = isi ? m.fi : m.fj;
    const long long nf = isi ? m.nfi : m.nfj;
    const long long cint = isi ? (long long)(b == 0 ? 0 : ni - 1) * nj + t : (long long)t * nj + (b == 2 ? 0 : nj - 1);
    FaceG<double> f{base[FS_AREA * nf + fidx], base[FS_N0 * nf + fidx], base[FS_N1 * nf + fidx]};
    if (b == 1 || b == 3) {
        f.n0 = -f.n0;
        f.n1 = -f.n1;
    }
    double qint[NV];
    for (int v = 0; v < NV; ++v) qint[v] = q[cint * NV + v];
    if (!ghost_value<double, LAW, BASIS>(sp, type, f, base[FS_CX * nf + fidx], base[FS_CY * nf + fidx], m.ccx[cint],
                                         m.ccy[cint], qint, qref, out))
        atomicOr(status, ST_BCNORMAL);
}

// max over cells of lambda / volume (spatial/spectralRadius.h:19-48; initial value
// numeric_limits::min(), :42): per work-group maxima, then the maximum of those (order-free)
__global__ void __launch_bounds__(256) k_specrad_partials(const double* __restrict__ res, const double* __restrict__ vol,
                                                          long long N, int nv, double* part)
{
    __shared__ double sm[256];
    double m = DBL_MIN;
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < N; c += (long long)gridDim.x * 256)
        m = fmax(m, res[c * (nv + 1) + nv] / vol[c]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(256) k_specrad_final(const double* part, int nparts, double* out)
{
    __shared__ double sm[256];
    double m = DBL_MIN;
    for (int p = threadIdx.x; p < nparts; p += 256) m = fmax(m, part[p]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0];
}

// Scratch of the allocation-free spectral radius (hc_spectral_radius_noalloc): the partial maxima
// of the first reduction pass (8 KiB); one copy per GPU, held until the process ends.  Host code
// never passes the array's name to a kernel: it takes the device address once per device
// (noalloc_setup) and passes that.
constexpr int NOALLOC_PARTS = 1024;
static __device__ __align__(256) double scratch[NOALLOC_PARTS];

// q1 = S( C(q0) + R (coef / vol) ) with mode 0 (eulerForwardUpdateGlobal, eulerForwardUpdate.h:38-92,
// coef = dt), or q1 = S( C(q0) + R (coef / lambda) ) with mode 1 (eulerForwardUpdateLocal, :117-168,
// coef = cfl, lambda of the given residual)
template <int LAW, int BASIS>
__global__ void __launch_bounds__(256) k_forward_update(int mode, double coef, Species<double> sp,
                                                        const double* __restrict__ vol, const double* __restrict__ res,
                                                        const double* __restrict__ q0, long long N,
                                                        double* __restrict__ q1, unsigned int* status)
{
    constexpr int NV = LawTraits<LAW>::nv;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const double* R = res + c * (NV + 1);
    const double fac = mode == 0 ? coef / vol[c] : coef / R[NV];
    double qc[NV], qn[NV];
    to_conserved<double, LAW, BASIS>(sp, q0 + c * NV, qc);
    for (int v = 0; v < NV; ++v) qc[v] = qc[v] + R[v] * fac;
    from_conserved<double, LAW, BASIS>(sp, qc, qn);
    bool fin = true;
    for (int v = 0; v < NV; ++v) {
        q1[c * NV + v] = qn[v];
        fin = fin && hc_isfinite(qn[v]);
    }
    if (!fin) atomicOr(status, ST_NONFINITE);
}

// legacy gradient (spatial/gradientCalc.h:16-313): per cell and direction the neighbour differences
// q_R - q_L of its edges in the reference's order (the i-edge to the lower neighbour, then to the
// upper one; the same along j), then the periodic wraps, whose loops stop one short of the last
// row / column (j < nj-1 for the i-wrap, i < ni-1 for the j-wrap).  dq [N][2][nv].
template <int LAW>
__global__ void __launch_bounds__(256) k_gradient(int ni, int nj, int per_i, int per_j, const double* __restrict__ q,
                                                  double* __restrict__ dq)
{
    constexpr int NV = LawTraits<LAW>::nv;
    const long long N = (long long)ni * nj;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const int i = (int)(c / nj), j = (int)(c % nj);
    for (int v = 0; v < NV; ++v) {
        const double qc = q[c * NV + v];
        double d0 = 0.0, d1 = 0.0;
        if (i > 0) d0 = d0 + (qc - q[(c - nj) * NV + v]);
        if (i < ni - 1) d0 = d0 + (q[(c + nj) * NV + v] - qc);
        if (j > 0) d1 = d1 + (qc - q[(c - 1) * NV + v]);
        if (j < nj - 1) d1 = d1 + (q[(c + 1) * NV + v] - qc);
        if (per_i && j < nj - 1 && (i == 0 || i == ni - 1))
            d0 = d0 + (q[(long long)j * NV + v] - q[((long long)(ni - 1) * nj + j) * NV + v]);
        if (per_j && i < ni - 1 && (j == 0 || j == nj - 1))
            d1 = d1 + (q[((long long)i * nj) * NV + v] - q[((long long)i * nj + nj - 1) * NV + v]);
        dq[(c * 2 + 0) * NV + v] = d0;
        dq[(c * 2 + 1) * NV + v] = d1;
    }
}

// Residual smoothing, the 1-D operators of OLD.fieldOperations/centralExplicitSmoothing.ipp:5-39
// and centralImplicitSmoothing.ipp:2-60 along the lines of one direction of the field
// r[N][stride] (values of variable v at r[c * stride + v]); dir 0 along i, 1 along j.
struct Line {
    int n;                 // points on the line
    long long step;        // cell-index step between neighbours on the line
    long long c;           // cell index of this point
    int k;                 // position on the line
};
__device__ inline Line line_of(long long c, int ni, int nj, int dir)
{
    const int i = (int)(c / nj), j = (int)(c % nj);
    if (dir == 0) return Line{ni, (long long)nj, c, i};
    return Line{nj, 1LL, c, j};
}

// explicit: r1[k] = (r0[k] + gamma (r0[k] - r0[k-1])) - gamma (r0[k+1] - r0[k]); on a periodic line
// the wrap edge (n-1, 0) adds to the end points last
__global__ void __launch_bounds__(256) k_smooth_explicit(int ni, int nj, int dir, int periodic, int nv, int stride,
                                                         double gamma, const double* __restrict__ r0,
                                                         double* __restrict__ r1)
{
    const long long N = (long long)ni * nj;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const Line L = line_of(c, ni, nj, dir);
    const long long first = c - (long long)L.k * L.step, last = first + (long long)(L.n - 1) * L.step;
    for (int v = 0; v < nv; ++v) {
        auto at = [&](long long cc) { return r0[cc * stride + v]; };
        const double x = at(c);
        double y = x;
        if (L.k > 0) y = y + gamma * (x - at(c - L.step));
        if (L.k < L.n - 1) y = y - gamma * (at(c + L.step) - x);
        if (periodic) {
            const double dr = gamma * (at(first) - at(last));
            if (L.k == L.n - 1) y = y - dr;
            if (L.k == 0) y = y + dr;
        }
        r1[c * stride + v] = y;
    }
}

// one weighted Jacobi sweep of (1 + 2 beta) r - beta (r[k-1] + r[k+1]) = r0:
//   s = (r0[k] + beta w[k-1]) + beta w[k+1], scaled by 1/(1 + 2 beta) (end points of a non-periodic
//   line: 1/(1 + beta)), w' = (1 - omega) w + omega s; the last sweep also multiplies by alpha
__global__ void __launch_bounds__(256) k_smooth_jacobi(int ni, int nj, int dir, int periodic, int nv, int stride,
                                                       double beta, double omega, double alpha, int last_sweep,
                                                       const double* __restrict__ r0, const double* __restrict__ w,
                                                       double* __restrict__ wn)
{
    const long long N = (long long)ni * nj;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const Line L = line_of(c, ni, nj, dir);
    const long long first = c - (long long)L.k * L.step, last = first + (long long)(L.n - 1) * L.step;
    const double s2 = 1. / (1. + 2. * beta), s1 = 1. / (1. + 1. * beta);
    for (int v = 0; v < nv; ++v) {
        double s = r0[c * stride + v];
        if (L.k > 0) s = s + beta * w[(c - L.step) * stride + v];
        if (L.k < L.n - 1) s = s + beta * w[(c + L.step) * stride + v];
        const bool end = L.k == 0 || L.k == L.n - 1;
        if (!end) s = s * s2;
        else if (periodic) {
            s = s + beta * w[(L.k == 0 ? last : first) * stride + v];
            s = s * s2;
        } else {
            s = s * s1;
        }
        double y = (1. - omega) * w[c * stride + v] + omega * s;
        if (last_sweep) y = y * alpha;
        wn[c * stride + v] = y;
    }
}

// ---- time step with residual smoothing (stepping of a single-case context) ----

// dt = cfl / max(part), dtslot[0] = dt, dtslot[1] += dt (the global step of stage 0, rungeKutta.h:133)
__global__ void __launch_bounds__(256) k_dt_from_max(const double* part, int nparts, double cfl, double* dtslot)
{
    __shared__ double sm[256];
    double m = DBL_MIN;
    for (int p = threadIdx.x; p < nparts; p += 256) m = fmax(m, part[p]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dt = cfl / sm[0];
        dtslot[0] = dt;
        dtslot[1] = dtslot[1] + dt;
    }
}

struct AlphaRow {
    double a[6];
};

// R_tot = sum_{k <= s} alpha[s][k] R_k in that order (rungeKutta.h:176-197); its lambda column holds
// stage 0's spectral-radius sum (the local time step's lambda)
template <int NV>
__global__ void __launch_bounds__(256) k_rk_total(int s, AlphaRow al, const double* __restrict__ R, long long N,
                                                  double* __restrict__ Rt)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    for (int v = 0; v < NV; ++v) {
        double t = 0.;
        for (int k = 0; k <= s; ++k) t = t + R[((long long)k * N + c) * (NV + 1) + v] * al.a[k];
        Rt[c * (NV + 1) + v] = t;
    }
    Rt[c * (NV + 1) + NV] = R[c * (NV + 1) + NV];
}

// q_out = S(C(q0) + R_tot fac), fac = (beta dt) / vol (global) or (beta cfl) / lambda_0 (local),
// on the device's [nv][N] state planes
template <int LAW, int BASIS>
__global__ void __launch_bounds__(256) k_update_planes(int global, double beta, double cfl, const double* dtslot,
                                                       Species<double> sp, const double* __restrict__ Rt,
                                                       const double* __restrict__ vol, const double* q0, long long N,
                                                       double* qout, unsigned int* status)
{
    constexpr int NV = LawTraits<LAW>::nv;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const double* R = Rt + c * (NV + 1);
    const double fac = global ? (beta * dtslot[0]) / vol[c] : (beta * cfl) / R[NV];
    double q[NV], qc[NV], qn[NV];
    for (int v = 0; v < NV; ++v) q[v] = q0[v * N + c];
    to_conserved<double, LAW, BASIS>(sp, q, qc);
    for (int v = 0; v < NV; ++v) qc[v] = qc[v] + R[v] * fac;
    from_conserved<double, LAW, BASIS>(sp, qc, qn);
    bool fin = true;
    for (int v = 0; v < NV; ++v) {
        qout[v * N + c] = qn[v];
        fin = fin && hc_isfinite(qn[v]);
    }
    if (!fin) atomicOr(status, ST_NONFINITE);
}

template <class F> hipError_t dispatch(int law, int basis, F&& f)
{
    if (law == HC_LAW_SCALAR) return f(std::integral_constant<int, 0>{}, std::integral_constant<int, 0>{});
    if (basis == HC_BASIS_PRIMITIVE) return f(std::integral_constant<int, 1>{}, std::integral_constant<int, 1>{});
    return f(std::integral_constant<int, 1>{}, std::integral_constant<int, 0>{});
}

unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

}   // namespace

using namespace hc_detail;

namespace {

int single_case(hc_context* ctx)
{
    int r = case_ready(ctx, 0);
    if (r) return r;
    if (ctx->cases.size() != 1) return fail(ctx, HC_ERR_STATE, "the spatial building blocks act on a single-case context");
    HC_HIP(hipSetDevice(ctx->device));
    return HC_OK;
}

MeshView<double> mesh_view(const hc_context* ctx)
{
    const CaseHost& c = ctx->cases[0];
    MeshView<double> m{};
    m.ni = c.ni;
    m.nj = c.nj;
    m.per_i = c.cs.bc[0] == HC_BC_PERIODIC;
    m.per_j = c.cs.bc[2] == HC_BC_PERIODIC;
    m.nfi = c.nfi;
    m.nfj = c.nfj;
    m.cvol = ctx->cvol;
    m.ccx = ctx->ccx;
    m.ccy = ctx->ccy;
    m.fi = ctx->fi;
    m.fj = ctx->fj;
    m.fi64 = ctx->fi;
    m.fj64 = ctx->fj;
    for (int d = 0; d < 2; ++d) {
        m.period_i[d] = c.cs.period_i[d];
        m.period_j[d] = c.cs.period_j[d];
    }
    return m;
}

Species<double> species_of(const hc_context* ctx)
{
    const hc_species& s = ctx->cases[0].cs.sp;
    return Species<double>{s.gamma, s.minf, s.lref, s.nu, s.pr, s.dt, s.R, s.gamma1};
}

// device staging buffer of the building blocks (grown on demand, released with the case)
int staging(hc_context* ctx, size_t ndoubles, double** out)
{
    if (ctx->blk_n < ndoubles) {
        HC_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->blk) (void)hipFree(ctx->blk);
        ctx->blk = nullptr;
        ctx->blk_n = 0;
        HC_HIP(hipMalloc(&ctx->blk, ndoubles * sizeof(double)));
        ctx->blk_n = ndoubles;
    }
    *out = ctx->blk;
    return HC_OK;
}

}   // namespace


namespace hc_detail {

constexpr int NOALLOC_MAX_DEVICES = 64;
static double* g_scratch[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// the device address of `scratch` on the current device (device), taken once per device
int noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return HC_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return HC_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return HC_ERR_HIP;
    g_scratch[device] = static_cast<double*>(ptr);
    return HC_OK;
}

// Residual smoothing of the flux part of r [N][stride] on the device (mode 0 explicit, 1 implicit,
// 2 explicit then implicit; each operator along i, then along j); b and c are scratch buffers of
// r's size; the result is left in r, the lambda column unchanged.  Enqueued on the context's stream.
int smooth_device(hc_context* ctx, int mode, double alpha, double* r, double* b, double* c, int stride)
{
    const CaseHost& cs = ctx->cases[0];
    const long long N = cs.N;
    const int nv = ctx->nv;
    const int per[2] = {cs.cs.bc[0] == HC_BC_PERIODIC, cs.cs.bc[2] == HC_BC_PERIODIC};
    const double beta = 0.25 * (alpha * alpha - 1.);
    const double gamma = 0.25 * ((1. + 4. * beta) / alpha - 1.);
    const double omega = 2. / 3.;
    const unsigned nb = blocks(N);
    const size_t bytes = (size_t)stride * N * sizeof(double);
    // the scratch buffers carry the lambda column along
    HC_HIP(hipMemcpyAsync(b, r, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HC_HIP(hipMemcpyAsync(c, r, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    const int first = mode == 1 ? 1 : 0, last = mode == 0 ? 0 : 1;
    for (int m = first; m <= last; ++m)
        for (int dir = 0; dir < 2; ++dir) {
            if (m == 0) {
                hipLaunchKernelGGL(k_smooth_explicit, dim3(nb), dim3(256), 0, ctx->stream, cs.ni, cs.nj, dir, per[dir],
                                   nv, stride, gamma, r, b);
                HC_HIP(hipGetLastError());
                HC_HIP(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            } else {
                // r holds r0 throughout; the iterates alternate r0 -> b -> c -> b -> c
                const double* w = r;
                for (int it = 0; it < 4; ++it) {
                    double* wn = (it % 2 == 0) ? b : c;
                    hipLaunchKernelGGL(k_smooth_jacobi, dim3(nb), dim3(256), 0, ctx->stream, cs.ni, cs.nj, dir,
                                       per[dir], nv, stride, beta, omega, alpha, it == 3 ? 1 : 0, r, w, wn);
                    HC_HIP(hipGetLastError());
                    w = wn;
                }
                HC_HIP(hipMemcpyAsync(r, c, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            }
        }
    return HC_OK;
}


// One time step of a single-case context with its stage residuals smoothed: per stage the residual
// of the stage state (the stage kernel in residual mode), the global time step at stage 0, the
// Runge-Kutta accumulation, the smoothing of the accumulated residual, and the forward-Euler update
// from the step-start state.  With the explicit operator at alpha = 1 (gamma = 0) this is the
// fused step bit for bit.  Enqueued on the context's stream.
int enqueue_step_smoothed(hc_context* ctx)
{
    const CaseHost& cs = ctx->cases[0];
    const long long N = cs.N;
    const int nv = ctx->nv, ns = ctx->cs.nstages;
    const size_t blk = (size_t)(nv + 1) * N;
    constexpr int NB = 1024;
    const size_t need = (size_t)(ns + 3) * blk + NB;
    if (ctx->sm_n < need) {
        HC_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->sm) (void)hipFree(ctx->sm);
        ctx->sm = nullptr;
        ctx->sm_n = 0;
        HC_HIP(hipMalloc(&ctx->sm, need * sizeof(double)));
        ctx->sm_n = need;
    }
    double* R = ctx->sm;
    double* Rt = R + (size_t)ns * blk;
    double* s1 = Rt + blk;
    double* s2 = s1 + blk;
    double* part = s2 + blk;
    const bool global = ctx->cs.dt_mode == HC_DT_GLOBAL;
    double* q0 = ctx->q[ctx->cur];
    double* bufA = ns == 1 ? ctx->q[1 - ctx->cur] : ctx->q[(ctx->cur + 1) % 3];
    double* bufB = ns == 1 ? bufA : ctx->q[(ctx->cur + 2) % 3];
    const Species<double> sp = species_of(ctx);
    const unsigned nb = blocks(N);
    const double* in = q0;
    for (int s = 0; s < ns; ++s) {
        double* out = ns == 1 ? bufA : (s == ns - 1 ? q0 : (s % 2 == 0 ? bufA : bufB));
        LaunchArgs<double> L = launch_args<double>(ctx);
        L.mode = MODE_RESIDUAL;
        L.q_in = in;
        L.q0 = in;
        L.q_out = out;   // not written in residual mode
        L.res_out = R + (size_t)s * blk;
        L.plan = ctx->plans[0];
        hipError_t e = launch_stage(ctx, L);
        if (e != hipSuccess) return hipfail(ctx, e, "residual kernel launch");
        if (s == 0 && global) {
            hipLaunchKernelGGL(k_specrad_partials, dim3(NB), dim3(256), 0, ctx->stream, R, ctx->cvol, N, nv, part);
            hipLaunchKernelGGL(k_dt_from_max, dim3(1), dim3(256), 0, ctx->stream, part, NB, ctx->cs.cfl, ctx->dtslot);
        }
        AlphaRow al{};
        for (int k = 0; k < 6; ++k) al.a[k] = ctx->cs.alpha[s][k];
        if (nv == 4) hipLaunchKernelGGL(k_rk_total<4>, dim3(nb), dim3(256), 0, ctx->stream, s, al, R, N, Rt);
        else hipLaunchKernelGGL(k_rk_total<3>, dim3(nb), dim3(256), 0, ctx->stream, s, al, R, N, Rt);
        HC_HIP(hipGetLastError());
        const int r = smooth_device(ctx, ctx->smooth_mode, ctx->smooth_alpha, Rt, s1, s2, nv + 1);
        if (r) return r;
        e = dispatch(ctx->cs.law, ctx->cs.basis, [&](auto Lw, auto B) {
            hipLaunchKernelGGL((k_update_planes<decltype(Lw)::value, decltype(B)::value>), dim3(nb), dim3(256), 0,
                               ctx->stream, global ? 1 : 0, ctx->cs.beta[s], ctx->cs.cfl, ctx->dtslot, sp, Rt, ctx->cvol,
                               q0, N, out, ctx->status);
            return hipGetLastError();
        });
        if (e != hipSuccess) return hipfail(ctx, e, "update kernel launch");
        in = out;
    }
    if (ns == 1) ctx->cur = 1 - ctx->cur;
    return HC_OK;
}

}   // namespace hc_detail

extern "C" {

int hc_xmetrics(hc_context* ctx, double* X)
{
    int r = single_case(ctx);
    if (r) return r;
    if (!X) return fail(ctx, HC_ERR_ARG, "null output");
    const std::vector<double>& x = ctx->cases[0].X;
    for (size_t k = 0; k < x.size(); ++k) X[k] = x[k];
    return HC_OK;
}

int hc_qmetrics(hc_context* ctx, const double* q, double* Q)
{
    int r = single_case(ctx);
    if (r) return r;
    if (!q || !Q) return fail(ctx, HC_ERR_ARG, "null array");
    const long long N = ctx->cases[0].N;
    const size_t nv = (size_t)ctx->nv;
    double* buf = nullptr;
    r = staging(ctx, 3 * nv * N, &buf);
    if (r) return r;
    double* dq = buf;
    double* dQ = buf + nv * N;
    HC_HIP(hipMemcpyAsync(dq, q, nv * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const MeshView<double> m = mesh_view(ctx);
    const hipError_t e = dispatch(ctx->cs.law, ctx->cs.basis, [&](auto L, auto) {
        hipLaunchKernelGGL((k_qmetrics<decltype(L)::value>), dim3(blocks(N)), dim3(256), 0, ctx->stream, m, dq, N, dQ);
        return hipGetLastError();
    });
    if (e != hipSuccess) return hipfail(ctx, e, "qmetrics kernel launch");
    HC_HIP(hipMemcpyAsync(Q, dQ, 2 * nv * N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_boundary_update(hc_context* ctx, const double* q, const double* bref, double* ghost)
{
    int r = single_case(ctx);
    if (r) return r;
    if (!q || !bref || !ghost) return fail(ctx, HC_ERR_ARG, "null array");
    const CaseHost& c = ctx->cases[0];
    const long long N = c.N;
    const size_t nv = (size_t)ctx->nv, rows = (size_t)c.brows;
    double* buf = nullptr;
    r = staging(ctx, nv * N + 2 * rows * nv, &buf);
    if (r) return r;
    double* dq = buf;
    double* dref = buf + nv * N;
    double* dghost = dref + rows * nv;
    HC_HIP(hipMemcpyAsync(dq, q, nv * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HC_HIP(hipMemcpyAsync(dref, bref, rows * nv * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    // rows of periodic boundaries are not written: keep the caller's values
    HC_HIP(hipMemcpyAsync(dghost, ghost, rows * nv * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const MeshView<double> m = mesh_view(ctx);
    const Species<double> sp = species_of(ctx);
    const int4 bc{c.cs.bc[0], c.cs.bc[1], c.cs.bc[2], c.cs.bc[3]};
    const hipError_t e = dispatch(ctx->cs.law, ctx->cs.basis, [&](auto L, auto B) {
        hipLaunchKernelGGL((k_boundary_update<decltype(L)::value, decltype(B)::value>), dim3(blocks((long long)rows)),
                           dim3(256), 0, ctx->stream, m, sp, bc, dq, dref, dghost, ctx->status);
        return hipGetLastError();
    });
    if (e != hipSuccess) return hipfail(ctx, e, "boundary update kernel launch");
    HC_HIP(hipMemcpyAsync(ghost, dghost, rows * nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return check_status(ctx);
}

int hc_spectral_radius(hc_context* ctx, const double* res, double* out)
{
    int r = single_case(ctx);
    if (r) return r;
    if (!res || !out) return fail(ctx, HC_ERR_ARG, "null array");
    const long long N = ctx->cases[0].N;
    const size_t nv = (size_t)ctx->nv;
    constexpr int NB = 1024;
    double* buf = nullptr;
    r = staging(ctx, (nv + 1) * N + NB + 1, &buf);
    if (r) return r;
    double* dres = buf;
    double* part = buf + (nv + 1) * N;
    HC_HIP(hipMemcpyAsync(dres, res, (nv + 1) * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_specrad_partials, dim3(NB), dim3(256), 0, ctx->stream, dres, ctx->cvol, N, (int)nv, part);
    hipLaunchKernelGGL(k_specrad_final, dim3(1), dim3(256), 0, ctx->stream, part, NB, part + NB);
    HC_HIP(hipGetLastError());
    HC_HIP(hipMemcpyAsync(out, part + NB, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_forward_update(hc_context* ctx, int mode, double coef, const double* res, const double* q0, double* q1)
{
    int r = single_case(ctx);
    if (r) return r;
    if (!res || !q0 || !q1) return fail(ctx, HC_ERR_ARG, "null array");
    if (mode != 0 && mode != 1) return fail(ctx, HC_ERR_ARG, "mode must be 0 (global) or 1 (local)");
    const long long N = ctx->cases[0].N;
    const size_t nv = (size_t)ctx->nv;
    double* buf = nullptr;
    r = staging(ctx, (nv + 1) * N + 2 * nv * N, &buf);
    if (r) return r;
    double* dres = buf;
    double* dq0 = buf + (nv + 1) * N;
    double* dq1 = dq0 + nv * N;
    HC_HIP(hipMemcpyAsync(dres, res, (nv + 1) * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HC_HIP(hipMemcpyAsync(dq0, q0, nv * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const Species<double> sp = species_of(ctx);
    const hipError_t e = dispatch(ctx->cs.law, ctx->cs.basis, [&](auto L, auto B) {
        hipLaunchKernelGGL((k_forward_update<decltype(L)::value, decltype(B)::value>), dim3(blocks(N)), dim3(256), 0,
                           ctx->stream, mode, coef, sp, ctx->cvol, dres, dq0, N, dq1, ctx->status);
        return hipGetLastError();
    });
    if (e != hipSuccess) return hipfail(ctx, e, "forward update kernel launch");
    HC_HIP(hipMemcpyAsync(q1, dq1, nv * N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return check_status(ctx);
}

int hc_gradient(hc_context* ctx, const double* q, double* dq)
{
    int r = single_case(ctx);
    if (r) return r;
    if (!q || !dq) return fail(ctx, HC_ERR_ARG, "null array");
    const CaseHost& c = ctx->cases[0];
    const long long N = c.N;
    const size_t nv = (size_t)ctx->nv;
    double* buf = nullptr;
    r = staging(ctx, 3 * nv * N, &buf);
    if (r) return r;
    double* dqin = buf;
    double* dout = buf + nv * N;
    HC_HIP(hipMemcpyAsync(dqin, q, nv * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const int per_i = c.cs.bc[0] == HC_BC_PERIODIC, per_j = c.cs.bc[2] == HC_BC_PERIODIC;
    const hipError_t e = dispatch(ctx->cs.law, ctx->cs.basis, [&](auto L, auto) {
        hipLaunchKernelGGL((k_gradient<decltype(L)::value>), dim3(blocks(N)), dim3(256), 0, ctx->stream, c.ni, c.nj,
                           per_i, per_j, dqin, dout);
        return hipGetLastError();
    });
    if (e != hipSuccess) return hipfail(ctx, e, "gradient kernel launch");
    HC_HIP(hipMemcpyAsync(dq, dout, 2 * nv * N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_smooth_residual(hc_context* ctx, int mode, double alpha, double* res)
{
    int r = single_case(ctx);
    if (r) return r;
    if (!res) return fail(ctx, HC_ERR_ARG, "null array");
    if (mode < 0 || mode > 2 || !(alpha >= 1.0)) return fail(ctx, HC_ERR_ARG, "smoothing mode 0..2 and alpha >= 1");
    const long long N = ctx->cases[0].N;
    const int stride = ctx->nv + 1;
    double* buf = nullptr;
    r = staging(ctx, 3 * (size_t)stride * N, &buf);
    if (r) return r;
    HC_HIP(hipMemcpyAsync(buf, res, (size_t)stride * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    r = smooth_device(ctx, mode, alpha, buf, buf + (size_t)stride * N, buf + 2 * (size_t)stride * N, stride);
    if (r) return r;
    HC_HIP(hipMemcpyAsync(res, buf, (size_t)stride * N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HC_HIP(hipStreamSynchronize(ctx->stream));
    return HC_OK;
}

int hc_set_residual_smoothing(hc_context* ctx, int mode, double alpha)
{
    if (!ctx) return HC_ERR_ARG;
    if (mode < -1 || mode > 2 || (mode >= 0 && !(alpha >= 1.0)))
        return fail(ctx, HC_ERR_ARG, "smoothing mode -1 (off) or 0..2 with alpha >= 1");
    if (mode >= 0) {
        int r = single_case(ctx);
        if (r) return r;
        if (ctx->cs.dt_mode == HC_DT_FIXED) return fail(ctx, HC_ERR_ARG, "residual smoothing needs a global or local time step");
    }
    drop_graph(ctx);
    ctx->smooth_mode = mode;
    ctx->smooth_alpha = mode >= 0 ? alpha : 1.0;
    return HC_OK;
}

// The spectral radius of a residual already on the device, without any allocation: the partial
// maxima live in the file-scope scratch of the GPU (every call overwrites the slots it reads, so
// nothing needs clearing); the result is written to the device address out_dev.  Asynchronous on
// the context's stream (hc_synchronize waits).  Calls that overlap on one GPU share its scratch
// and race.
int hc_spectral_radius_noalloc(hc_context* ctx, const double* res_dev, double* out_dev)
{
    int r = single_case(ctx);
    if (r) return r;
    if (!res_dev || !out_dev) return fail(ctx, HC_ERR_ARG, "null device pointer");
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) return fail(ctx, HC_ERR_ARG, "device index out of range");
    double* part = g_scratch[ctx->device];
    if (!part) return fail(ctx, HC_ERR_STATE, "scratch address of the device not set up (hc_create)");
    HC_HIP(hipSetDevice(ctx->device));
    const long long N = ctx->cases[0].N;
    const long long nb = (N + 255) / 256;
    const int nparts = (int)(nb < NOALLOC_PARTS ? nb : NOALLOC_PARTS);   // grid-stride: any N fits
    if (nparts < 1 || nparts > NOALLOC_PARTS) return fail(ctx, HC_ERR_ARG, "partial count beyond the scratch");
    hipLaunchKernelGGL(k_specrad_partials, dim3(nparts), dim3(256), 0, ctx->stream, res_dev, ctx->cvol, N, ctx->nv,
                       part);
    hipLaunchKernelGGL(k_specrad_final, dim3(1), dim3(256), 0, ctx->stream, part, nparts, out_dev);
    HC_HIP(hipGetLastError());
    return HC_OK;
}

}   // extern "C"
