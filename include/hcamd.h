/*
 * hcamd.h -- C-ABI of the MI355X (gfx950) structured-mesh finite-volume hot path:
 * spatial residual (MUSCL + limiter + flux functor, LSQ gradients, boundary
 * conditions) and explicit Runge-Kutta time stepping, executed by hand-written
 * HIP kernels on one GPU.
 *
 * The reference (lacrymose/hyperboliConvergence) exposes this path as a
 * header-only C++20 template surface; its C++ callers keep that surface
 * (hyperboliconvergence_amd/cpp/, "hcamd/hcamd.hpp"), and the templates forward
 * here.  Each entry point names the reference interface it replaces
 * (paths relative to the reference's cpp/inc/).
 *
 * Conventions
 *  - int status return: 0 ok; HC_ERR_* < 0 on invalid arguments or HIP errors;
 *    HC_ERR_STATUS when the device raised a status bit (hc_get_status).
 *  - Host arrays are borrowed for the duration of the call and are row-major
 *    with the last index fastest, like the reference's par::Array
 *    (parallalg/array.h:267-349):
 *      nodes [ni+1][nj+1][2], q [ni][nj][nvar], res [ni][nj][nvar+1],
 *      bref  [2*nj + 2*ni][nvar] (boundary ids 0,1 along j, then 2,3 along i).
 *  - One context per host thread; a context is not thread-safe; calls on a
 *    context are serialised on its HIP stream and are synchronous unless noted.
 */
#ifndef HCAMD_H
#define HCAMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define HC_OK 0
#define HC_ERR_ARG (-1)
#define HC_ERR_HIP (-2)
#define HC_ERR_STATE (-3)
#define HC_ERR_STATUS (-4)

/* identifiers (same numbering as the oracle and the C++ drop-in layer) */
enum { HC_LAW_SCALAR = 0, HC_LAW_EULER = 1 };
enum { HC_BASIS_CONSERVED = 0, HC_BASIS_PRIMITIVE = 1 };
/* HC_FLUX_ROE_FACEFRAME: Roe with the Roe-average velocity in the face frame (no reference counterpart);
   HC_FLUX_ROE_UNPREC_WS: RoeUnprecWS (conservationLaws/euler/euler.h:392-405, euler/fluxes/roeUnprecWS.ipp) */
enum { HC_FLUX_CENTRAL = 0, HC_FLUX_RUSANOV = 1, HC_FLUX_ROE = 2, HC_FLUX_SLAU = 3, HC_FLUX_AUSMPUP = 4,
       HC_FLUX_ROE_FACEFRAME = 5, HC_FLUX_ROE_UNPREC_WS = 6 };
enum { HC_SCALE_CONVECTIVE = 0, HC_SCALE_ACOUSTIC = 1, HC_SCALE_ADAPTIVE = 2 };
enum { HC_LIM_NOLIMIT1 = 0, HC_LIM_NOLIMIT2 = 1, HC_LIM_NOLIMIT3 = 2, HC_LIM_VANALBADA2 = 3,
       HC_LIM_MINMOD2 = 4, HC_LIM_SUPERBEE2 = 5, HC_LIM_MC2 = 6, HC_LIM_VANLEER2 = 7,
       HC_LIM_CADA3 = 8, HC_LIM_SMOOTHVANALBADA = 9 };
enum { HC_BC_PERIODIC = 0, HC_BC_RIEMANN = 1, HC_BC_ENTROPY = 2, HC_BC_FIXED = 3, HC_BC_WALL = 4 };
/* HC_DT_FIXED: hc_case_desc.cfl holds the time step itself; no spectral-radius reduction
   (the differentiable form of the time loop, used by the adjoint) */
enum { HC_DT_GLOBAL = 0, HC_DT_LOCAL = 1, HC_DT_FIXED = 2 };
enum { HC_PRECISION_FP64 = 0, HC_PRECISION_MIXED = 1 };

/* status bits (hc_get_status) */
#define HC_STATUS_NONFINITE 1u
#define HC_STATUS_NONPHYSICAL 2u
#define HC_STATUS_BCNORMAL 4u

/* Species<Euler,Real> (conservationLaws/euler/euler.h:48-60) */
typedef struct {
    double gamma, minf, lref, nu, pr, dt, R, gamma1;
} hc_species;

/*
 * Case descriptor: the runtime form of the reference's template arguments of
 * make_muscl_flux<Law>(Limiter{}, Flux{}) (spatial/muscl.h:10-60), the boundary
 * condition tuple (spatial/boundary/boundaryCondition.h:57-136), the Runge-Kutta
 * table (ode.h:87-105), UnsteadyTimeControls::cfl (controls.h:11-19) and Species.
 * period_i / period_j: translation from a cell to its periodic image across
 * boundary 1 / 3 (zero for meshes that close on themselves, such as an O-grid).
 */
typedef struct {
    int law, basis, flux, vscale, pscale, limiter;
    int bc[4];
    int nstages, dt_mode;
    double alpha[6][6];
    double beta[6];
    double cfl;
    double period_i[2];
    double period_j[2];
    hc_species sp;
} hc_case_desc;

typedef struct hc_context hc_context;

/* Create / destroy a context on a HIP device (owns a stream and device buffers); destroying a
   null context does nothing.  Every other entry point that takes a context returns HC_ERR_ARG
   for a null one (hc_last_error returns the message "null context"). */
int hc_create(hc_context** ctx, int device);
int hc_destroy(hc_context* ctx);
const char* hc_last_error(const hc_context* ctx);

/*
 * Mesh nodes of a structured 2-D mesh; replaces Mesh<2,Real> construction and
 * dual() (mesh/mesh.h:19-57, mesh/dual.ipp:42-83).  Geometry is derived on
 * the next hc_set_case.
 */
int hc_set_mesh(hc_context* ctx, int ni, int nj, const double* nodes);

/*
 * Flux functor, limiter, boundary conditions, Runge-Kutta table, CFL and
 * species; builds the one-time geometry (cells, faces, LSQ metrics) and
 * replaces make_muscl_flux (spatial/muscl.h:10-87), xmetrics
 * (spatial/lsqMetrics.h:20-62) and the BC tuple.
 */
int hc_set_case(hc_context* ctx, const hc_case_desc* desc);

/* FP64 throughout (default) or FP32 face arithmetic with FP64 state, RK and metrics. */
int hc_set_precision(hc_context* ctx, int precision);

/* Boundary reference values, SolutionField::boundary[b](k,1) (solutionField/solutionField.h:56-91). */
int hc_set_boundary_refs(hc_context* ctx, const double* bref);

/* Solution interior in the solution basis, SolutionField::interior. */
int hc_set_state(hc_context* ctx, const double* q);
int hc_get_state(hc_context* ctx, double* q);

/*
 * Spatial residual of the current state, FluxResult per cell {flux[nvar], lambda}:
 * boundaryUpdate + qmetrics + residualCalc (spatial/boundary/boundaryUpdate.h:28-48,
 * spatial/lsqMetrics.h:181-213, spatial/residualCalc.h:45-69).
 */
int hc_residual(hc_context* ctx, double* res);

/*
 * Advance nsteps time steps: the body of integrate() (timestepping/rungeKutta.h:44-167)
 * with rungeKuttaAccumulation (:176-197), spectralRadius (spatial/spectralRadius.h:24-48)
 * and eulerForwardUpdateGlobal / Local (spatial/eulerForwardUpdate.h:38-168).
 * Asynchronous with respect to the host (hc_synchronize or any getter waits).
 */
int hc_step(hc_context* ctx, int nsteps);

/*
 * Steady-state marching: the parameters of SteadyTimeControls (controls.h:31-48; at the
 * reference's HEAD no integrate() consumes them).  Steps in chunks of check_every (graphs as
 * hc_step) until the L2 norm of the residual's third variable (mass for Euler, the transported
 * scalar for advection; see hc_norms) has dropped to residual_drop times its value at the
 * start, or max_steps are taken.  Single-case contexts.  On return *steps_done holds the steps
 * taken and *drop the last ratio (1 if the start residual is zero).
 */
int hc_step_steady(hc_context* ctx, long long max_steps, double residual_drop, int check_every,
                   long long* steps_done, double* drop);

/* Sum of the global time steps taken (the reference's t, rungeKutta.h:96,163). */
int hc_get_time(hc_context* ctx, double* t);

/*
 * L2 (root mean square) and Linf norms per variable of R/vol of the current
 * state's residual, accumulated in FP64.  No counterpart at the reference's
 * HEAD (OLD.fieldOperations/fieldOperations.h:38-42 declares a residual()).
 */
int hc_norms(hc_context* ctx, double* l2, double* linf);

/*
 * hc_step_norms: nsteps steps (>= 1) as hc_step, and the norms (as hc_norms) of the residual of
 * the state the last step started from, which that step's first stage computes anyway: the stage
 * kernel also stores it and a reduction pass folds it, instead of a separate residual evaluation.
 * Single-case contexts.  The state afterwards is hc_step's; the norms equal hc_step(nsteps - 1)
 * followed by hc_norms.  No counterpart in the reference.
 */
int hc_step_norms(hc_context* ctx, int nsteps, double* l2, double* linf);

/* OR of the status bits raised by the device since the last call (cleared on read). */
int hc_get_status(hc_context* ctx, unsigned int* status);

int hc_synchronize(hc_context* ctx);

/*
 * Capture the kernels of k consecutive steps in a HIP graph and replay it in
 * hc_step (k = 0 disables graphs).  Replaces nothing in the reference (its
 * integrate() loop runs on the host).
 */
int hc_set_graph_steps(hc_context* ctx, int k);

/*
 * Kernel timing on the context's stream with HIP events.  mode 1: events around every
 * launch (graphs off); mode 2: one event pair around each hc_step call (graphs on), the
 * interval counted against the stage kernels it launched.  hc_get_profile returns the sums
 * {stage kernel ms, stage-kernel launches, update kernel ms, update launches} since the last
 * hc_set_profiling.  Used for the roofline report.
 */
int hc_set_profiling(hc_context* ctx, int enable);
int hc_get_profile(hc_context* ctx, double* out4);

/* Device pointer of the current state ([nvar][ni][nj] FP64, valid until the next call). */
int hc_device_state(hc_context* ctx, double** dptr);

/*
 * STREAM triad a[i] = b[i] + s*c[i] over three device arrays of n doubles (n >= 2), repeated
 * reps times with 16-byte loads and stores; returns the attainable HBM bandwidth in GB/s
 * (24 bytes per double of the even part of n per repetition).  Then, once, a[i] = s*b[i] over the
 * same even part with 8-byte accesses (the stage kernels' width).  Used as the attainable-bandwidth
 * denominator and to calibrate memory counters at both access widths.  No counterpart in the
 * reference.
 */
int hc_stream_triad(hc_context* ctx, long long n, int reps, double* gbps);

/*
 * Batches.  A context can hold several cases (meshes with their own boundary types, periodic
 * translations, species and boundary references) that share the law, basis, flux, limiter,
 * scalings, Runge-Kutta table, time-step mode and CFL: every stage launch then covers the
 * tiles of all cases (one tile table), and hc_step advances all of them.  The reference runs
 * one integrate() per case (timestepping/rungeKutta.h:44-167); these entry points have no
 * reference counterpart.  Sequence: hc_batch_begin, hc_batch_add_case per case (geometry is
 * built here), hc_batch_end (device layout), then per case hc_batch_set_boundary_refs and
 * hc_batch_set_state.  The single-case entry points above act on case 0.
 */
int hc_batch_begin(hc_context* ctx);
int hc_batch_add_case(hc_context* ctx, int ni, int nj, const double* nodes, const hc_case_desc* desc,
                      int* case_id);
int hc_batch_end(hc_context* ctx);
int hc_batch_size(hc_context* ctx, int* ncases);
int hc_batch_set_boundary_refs(hc_context* ctx, int case_id, const double* bref);
int hc_batch_set_state(hc_context* ctx, int case_id, const double* q);
int hc_batch_get_state(hc_context* ctx, int case_id, double* q);
int hc_batch_residual(hc_context* ctx, int case_id, double* res);
int hc_batch_norms(hc_context* ctx, int case_id, double* l2, double* linf);
int hc_batch_get_time(hc_context* ctx, int case_id, double* t);

/*
 * Tile configuration of the stage kernels (shape, and whether the centre differences are
 * formed from cell centres in LDS): config -1 (default) times every configuration on the first
 * steps of the context (all give bit-identical results) and keeps the fastest; 0 .. 2 fixes one.
 * hc_get_tiles returns the configuration in use and, after timing, the milliseconds per step of
 * each (HC_NTILECFG values).  No counterpart in the reference.
 */
#define HC_NTILECFG 3
int hc_set_tiles(hc_context* ctx, int config);
int hc_get_tiles(hc_context* ctx, int* config, double* ms_per_step);

/*
 * Uniform geometry: when every case of the context is a periodic box whose face geometry and cell
 * volumes are bitwise the same for all faces of a kind and all cells (a uniform Cartesian grid),
 * the stage kernels read one set of face slots per kind and one volume per case instead of one
 * per face and cell; results are bit-identical.  enable = 1 (default) uses those kernels whenever
 * the mesh allows it, 0 always uses the general ones.  hc_get_uniform_geometry reports whether the
 * uniform kernels are in use.  No counterpart in the reference.
 */
int hc_set_uniform_geometry(hc_context* ctx, int enable);
int hc_get_uniform_geometry(hc_context* ctx, int* active);

/*
 * Shared geometry and tile order of a batch: cases whose cell and face geometry is bitwise equal
 * (the same mesh and boundary kinds, as in a sweep of flow conditions over a few meshes) share one
 * device copy of it.  interleave = 1 (default) orders a step's work-groups so that the cases of
 * such a group run tile by tile together (their face slots are fetched once and then read from
 * L2); 0 runs the cases one after the other.  Results are bit-identical either way.
 * hc_get_geometry_blocks returns the number of geometry copies held on the device.  No
 * counterpart in the reference (one mesh per integrate() call there).
 */
int hc_set_tile_order(hc_context* ctx, int interleave);
int hc_get_geometry_blocks(hc_context* ctx, int* nblocks);

/*
 * Derivatives (single-case context, FP64; Euler with every flux in the conserved basis, AUSM+up
 * also in the primitive basis).  No counterpart in the reference.
 * Parameters: the state q, minf of the species (AUSM+up / SLAU / RoeUnprecWS scaling) and the boundary
 * reference states bref ([rows][nvar] as hc_set_boundary_refs).
 *
 * hc_tangent_residual: R(q) of the current state ([N][nvar+1], as hc_residual) and
 *   Rdot = dR/dq qdot + dR/dbref bref_dot + dR/dminf minf_dot ([N][nvar]); bref_dot may be null.
 * hc_adjoint_residual: qbar = (dR/dq)^T ybar ([N][nvar]), bref_bar = (dR/dbref)^T ybar and
 *   minf_bar = (dR/dminf)^T ybar, for ybar [N][nvar].
 *
 * Trajectories (conserved basis, HC_DT_FIXED): hc_adjoint_forward advances the current state
 * nsteps steps and keeps a checkpoint every checkpoint_every steps; hc_adjoint_backward then
 * maps the adjoint of the final state to the adjoint of the initial state and the adjoints of
 * bref and minf accumulated over all stages, recomputing each segment's states from its
 * checkpoint.  hc_tangent_steps advances the state nsteps steps together with its tangent
 * (initial qdot0, parameter tangents bref_dot, minf_dot) and returns the final tangent.
 */
int hc_tangent_residual(hc_context* ctx, const double* qdot, const double* bref_dot, double minf_dot, double* res,
                        double* res_dot);
int hc_adjoint_residual(hc_context* ctx, const double* ybar, double* qbar, double* bref_bar, double* minf_bar);
int hc_adjoint_forward(hc_context* ctx, int nsteps, int checkpoint_every);
int hc_adjoint_backward(hc_context* ctx, const double* qbar_final, double* qbar_initial, double* bref_bar,
                        double* minf_bar);
int hc_tangent_steps(hc_context* ctx, int nsteps, const double* qdot0, const double* bref_dot, double minf_dot,
                     double* qdot_final);

/*
 * Device-array variants for callers whose data is already in device memory of the HIP runtime the
 * library uses (for instance PyTorch tensors on the GPU, in a process where torch's HIP runtime is
 * the one loaded): the same layouts and meaning as hc_set_state, hc_get_state,
 * hc_set_boundary_refs (solutionField/solutionField.h:10-140, the SolutionField's interior and
 * boundary arrays) and hc_adjoint_backward, for the context's single case; minf_bar stays a host
 * scalar.  The caller's writes to the inputs must be complete before the call (its stream
 * synchronised); the calls return once the outputs are written.  A pointer that is not device
 * memory of that runtime is rejected with HC_ERR_ARG.
 */
int hc_set_state_device(hc_context* ctx, const double* q_dev);
int hc_get_state_device(hc_context* ctx, double* q_dev);
int hc_set_boundary_refs_device(hc_context* ctx, const double* bref_dev);
int hc_adjoint_backward_device(hc_context* ctx, const double* qbar_final_dev, double* qbar_initial_dev,
                               double* bref_bar_dev, double* minf_bar);

/*
 * Spatial building blocks of a single-case context, for callers that run the time loop
 * themselves on host arrays (the reference's integrate() body, timestepping/rungeKutta.h:97-164);
 * hc_step runs the same arithmetic fused.  Arrays as above; the geometry, boundary types, basis
 * and species are the context's.
 *
 * hc_xmetrics: the LSQ spatial metric per cell X [N][3] = {xx, xy, yy} (xmetrics,
 *   spatial/lsqMetrics.h:20-62 with boundary_xmetrics :104-163; mirror points on non-periodic
 *   sides, periodic images across periodic ones).
 * hc_qmetrics: the LSQ solution metric Q [N][nvar][2] of the state q (qmetrics,
 *   spatial/lsqMetrics.h:181-213, lsq/qmetric.ipp:48-70).
 * hc_boundary_update: ghost values [2nj+2ni][nvar] of the non-periodic boundaries from q and the
 *   reference values bref (boundaryUpdate, spatial/boundary/boundaryUpdate.h:28-48, 137-266); rows
 *   of periodic boundaries are left as the caller passed them.
 * hc_spectral_radius: max over cells of lambda / volume of a residual [N][nvar+1] (spectralRadius,
 *   spatial/spectralRadius.h:19-48, initial value DBL_MIN).
 * hc_forward_update: q1 = S(C(q0) + R coef / vol) with mode 0 (eulerForwardUpdateGlobal,
 *   spatial/eulerForwardUpdate.h:38-92, coef = dt) or q1 = S(C(q0) + R coef / lambda) with mode 1
 *   (eulerForwardUpdateLocal, :117-168, coef = cfl, lambda of the residual passed).
 */
int hc_xmetrics(hc_context* ctx, double* X);
int hc_qmetrics(hc_context* ctx, const double* q, double* Q);
int hc_boundary_update(hc_context* ctx, const double* q, const double* bref, double* ghost);
int hc_spectral_radius(hc_context* ctx, const double* res, double* out);
int hc_forward_update(hc_context* ctx, int mode, double coef, const double* res, const double* q0, double* q1);

/*
 * hc_gradient: the legacy neighbour-difference gradient dq [N][2][nvar] of the state q
 *   (gradientCalc, spatial/gradientCalc.h:16-313): per direction the sum of q_R - q_L over the
 *   cell's edges, the periodic wraps stopping one short of the last row / column like the
 *   reference's loops (:275, :296).
 * hc_smooth_residual: residual smoothing of res [N][nvar+1] in place (the flux part), the product
 *   of the reference's 1-D operators along i then along j (OLD.fieldOperations/
 *   centralExplicitSmoothing.ipp:5-39, centralImplicitSmoothing.ipp:2-60): mode 0 explicit
 *   (gamma = ((1 + 4 beta)/alpha - 1)/4), 1 implicit (4 weighted Jacobi sweeps, beta =
 *   (alpha^2 - 1)/4, times alpha), 2 explicit then implicit; alpha >= 1 is the time-step factor.
 *   A direction is periodic when its boundaries are.
 */
int hc_gradient(hc_context* ctx, const double* q, double* dq);
int hc_smooth_residual(hc_context* ctx, int mode, double alpha, double* res);

/*
 * Residual smoothing in the time loop (SteadyTimeControls::use_residual_smoothing and
 * smoothing_factor, controls.h:31-48, which no integrate() consumes at the reference's HEAD):
 * mode 0..2 and alpha as hc_smooth_residual, applied to every stage's accumulated residual
 * before its update in hc_step / hc_step_steady (an unfused step: residual, accumulation,
 * smoothing and update as separate kernels); mode -1 switches it off.  Single-case contexts with
 * a global or local time step.
 */
int hc_set_residual_smoothing(hc_context* ctx, int mode, double alpha);

/*
 * Allocation-free variant of hc_spectral_radius for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid allocation); it has no
 * counterpart in the reference.  res_dev: a device array [N][nvar+1] of the context's (single)
 * case; out_dev: device address of one double that receives max over cells of lambda / volume.
 * The reduction's partial maxima live in a fixed per-GPU scratch of 1024 doubles that the library
 * holds until the process ends (any N fits: the first pass strides over the cells with at most
 * 1024 work-groups); the call allocates nothing and is asynchronous on the context's stream
 * (hc_synchronize waits).  Calls that overlap on one GPU share that scratch and race.
 */
int hc_spectral_radius_noalloc(hc_context* ctx, const double* res_dev, double* out_dev);

/* Library version string. */
const char* hc_version(void);

#ifdef __cplusplus
}
#endif
#endif
