// hc_context.hpp -- the context behind the C-ABI (include/hcamd.h) and the internal helpers
// shared by the translation units of the library (hc_capi.cpp, hc_adjoint.cpp).
#pragma once

#include "hcamd.h"

#include "hc_geom.hpp"
#include "hc_launch.hpp"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

using hc::CaseRec;
using hc::GeomOutput;
using hc::NTILECFG;
using hc::RkPlan;
using hc::UpdChunk;

struct AdjState;   // buffers of the tangent / adjoint entry points (hc_adjoint.cpp)

namespace hc_detail {

// host-side record of one case of a context
struct CaseHost {
    int ni = 0, nj = 0;
    hc_case_desc cs{};
    GeomOutput g;                  // host geometry until it is uploaded
    std::vector<double> X;         // LSQ spatial metric per cell {xx, xy, yy}, kept for hc_xmetrics
    long long N = 0, nfi = 0, nfj = 0;
    long long cell_off = 0, bref_off = 0;
    // geometry block: the first case whose geometry is bitwise equal to this one's, and the
    // block's offsets into the packed geometry arrays (cells, i-faces, j-faces)
    int geo = 0;
    long long geo_cell_off = 0, geo_fi_off = 0, geo_fj_off = 0;
    int brows = 0;                 // boundary reference rows 2 nj + 2 ni
    bool have_state = false;
    bool uniform = false;          // uniform geometry (hc_geom.hpp uniform_geometry)
    int tiles_j[NTILECFG] = {}, tile0[NTILECFG] = {}, ntiles[NTILECFG] = {};
};

}   // namespace hc_detail

struct hc_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // single-case staging (hc_set_mesh)
    int ni = 0, nj = 0;
    std::vector<double> nodes;
    bool have_mesh = false;
    // cases
    bool building = false;         // between hc_batch_begin and hc_batch_end
    std::vector<hc_detail::CaseHost> cases;
    bool have_case = false;        // device layout built
    hc_case_desc cs{};             // parameters shared by the cases (case 0's descriptor)
    int nv = 0;
    int precision = HC_PRECISION_FP64;
    std::vector<RkPlan> plans;
    int nslots = 0;
    long long Ntot = 0, nfi_tot = 0, nfj_tot = 0, brows_tot = 0;
    // device buffers
    double *cvol = nullptr, *ccx = nullptr, *ccy = nullptr, *fi = nullptr, *fj = nullptr;
    float *fi32 = nullptr, *fj32 = nullptr;
    double* q[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    double *acc = nullptr, *tmp = nullptr, *lam0 = nullptr, *bref = nullptr, *res = nullptr, *part = nullptr,
           *nout = nullptr;
    unsigned long long* maxslot = nullptr;   // MAXSPREAD slots per case, MAXSTRIDE words apart
    double* dtslot = nullptr;                // two per case
    unsigned int* status = nullptr;
    CaseRec<double>* rec64[NTILECFG] = {};
    CaseRec<float>* rec32[NTILECFG] = {};
    hc::TileRef* tiles[NTILECFG] = {};   // tile table per tile shape: step order, then each case's tiles
    int ntiles[NTILECFG] = {};           // tiles of a step (the step-order part of the table)
    bool interleave = true;              // step order interleaves cases that share geometry (hc_set_tile_order)
    int ngeo = 0;                        // geometry blocks held on the device
    UpdChunk* chunks = nullptr;
    int nchunks = 0;
    // tile shape: fixed index, or -1 to time every shape on the first steps and keep the fastest
    int tile_mode = -1;
    int tile_cfg = 0;
    bool uni_ok = false;           // every case has uniform geometry
    bool uni_enable = true;        // use the uniform-geometry kernels when uni_ok (hc_set_uniform_geometry)
    bool tuned = false;
    double tune_ms[NTILECFG] = {};
    // graphs
    int graph_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_cur = -1, graph_cfg = -1;
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
    // tangent / adjoint buffers, created on first use
    AdjState* adj = nullptr;
    // device staging buffer of the spatial building blocks (hc_qmetrics ... hc_forward_update)
    double* blk = nullptr;
    size_t blk_n = 0;
    // residual smoothing of the stepping (hc_set_residual_smoothing): mode -1 off, its buffers
    bool keep_res = false;         // enqueue_step: stage 0 also stores its residual in res (hc_step_norms)
    int smooth_mode = -1;
    double smooth_alpha = 1.0;
    double* sm = nullptr;
    size_t sm_n = 0;
};


namespace hc_detail {

int fail(hc_context* c, int code, const std::string& msg);
int hipfail(hc_context* c, hipError_t e, const char* where);
int ready(hc_context* ctx);
int case_ready(hc_context* ctx, int k);
int check_status(hc_context* ctx);
int enqueue_step(hc_context* ctx);
hipError_t launch_stage(hc_context* ctx, const hc::LaunchArgs<double>& L);
template <class RF> hc::LaunchArgs<RF> launch_args(hc_context* ctx);
void release_adjoint(hc_context* ctx);
int smooth_device(hc_context* ctx, int mode, double alpha, double* r, double* b, double* c, int stride);
int enqueue_step_smoothed(hc_context* ctx);
int noalloc_setup(int device);
bool device_pointer(const void* p);
int aos_to_soa_device(hc_context* ctx, double* soa, const double* aos_dev, long long N, int nv);
int soa_to_aos_device(hc_context* ctx, double* aos_dev, const double* soa, long long N, int nv);
void drop_graph(hc_context* ctx);

}   // namespace hc_detail

#define HC_HIP(call)                                                      \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return hc_detail::hipfail(ctx, e_, #call);  \
    } while (0)
