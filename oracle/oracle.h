/*
 * oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the
 * reference's structured-mesh finite-volume residual and explicit Runge-Kutta
 * path (lacrymose/hyperboliConvergence, cpp/inc/...).  It is the checker for
 * the HIP product path and the timed CPU baseline; nothing in the product
 * links, loads or calls it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Conventions (all arrays are row-major, last index fastest, like the
 * reference's par::Array, cpp/inc/parallalg/array.h):
 *   nodes  [ni+1][nj+1][2]            primal mesh points
 *   cells  [ni][nj][3]                {volume, centre x, centre y}
 *   q      [ni][nj][nvar]             solution variables (solution basis)
 *   bref   [2*nj + 2*ni][nvar]        boundary reference values, boundary ids
 *                                     0,1 (nj entries each) then 2,3 (ni each),
 *                                     i.e. SolutionField::boundary[b](k,1)
 *   res    [ni][nj][nvar+1]           FluxResult {flux[nvar], lambda}
 */
#ifndef HC_ORACLE_H
#define HC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double gamma, minf, lref, nu, pr, dt, R, gamma1;
} orc_species;

/* Identifiers shared with the product's C-ABI (same numbering). */
enum { ORC_LAW_SCALAR = 0, ORC_LAW_EULER = 1 };
enum { ORC_BASIS_CONSERVED = 0, ORC_BASIS_PRIMITIVE = 1 };
enum { ORC_FLUX_CENTRAL = 0, ORC_FLUX_RUSANOV = 1, ORC_FLUX_ROE = 2,
       ORC_FLUX_SLAU = 3, ORC_FLUX_AUSMPUP = 4, ORC_FLUX_ROE_FACEFRAME = 5,
       ORC_FLUX_ROE_UNPREC_WS = 6 };
enum { ORC_SCALE_CONVECTIVE = 0, ORC_SCALE_ACOUSTIC = 1, ORC_SCALE_ADAPTIVE = 2 };
enum { ORC_LIM_NOLIMIT1 = 0, ORC_LIM_NOLIMIT2 = 1, ORC_LIM_NOLIMIT3 = 2,
       ORC_LIM_VANALBADA2 = 3, ORC_LIM_MINMOD2 = 4, ORC_LIM_SUPERBEE2 = 5,
       ORC_LIM_MC2 = 6, ORC_LIM_VANLEER2 = 7, ORC_LIM_CADA3 = 8,
       ORC_LIM_SMOOTHVANALBADA = 9 };
enum { ORC_BC_PERIODIC = 0, ORC_BC_RIEMANN = 1, ORC_BC_ENTROPY = 2,
       ORC_BC_FIXED = 3, ORC_BC_WALL = 4 };
/* ORC_DT_FIXED: cfl holds the time step itself (no spectral-radius reduction) */
enum { ORC_DT_GLOBAL = 0, ORC_DT_LOCAL = 1, ORC_DT_FIXED = 2 };

typedef struct {
    int law, basis, flux, vscale, pscale, limiter;
    int bc[4];
    int nstages, dt_mode;
    double alpha[6][6];
    double beta[6];
    double cfl;
    double period_i[2];   /* translation from a cell to its periodic image across boundary 1 */
    double period_j[2];   /* ... across boundary 3 */
    orc_species sp;
} orc_case;

typedef struct {
    double area, cx, cy, nx, ny, tx, ty;   /* geom::Surface: area, centre, metric[0], metric[1] */
} orc_face;

int  orc_nvar(int law);

/* geometry / mesh (geometry/surface.ipp, volume.ipp, mesh/dual.ipp, mesh/generate/twoD.h) */
void orc_surface(const double* p0, const double* p1, orc_face* out);
void orc_volume(const double* p0, const double* p1, const double* p2, const double* p3, double* out3);
void orc_dual(int ni, int nj, const double* nodes, double* cells);
void orc_mesh_linspace(int ni, int nj, double lox, double hix, double loy, double hiy, double* nodes);
void orc_mesh_cylinder(int nr, int nth, double rc, double rb, double* nodes);
double orc_ogrid_ratio(int nr, double rc, double rb, double dr0);
void orc_mesh_ogrid(int nr, int nth, double rc, double rb, double dr0, double* nodes);

/* physics point functions */
orc_species orc_air_species(void);
void orc_set2state(const orc_case* cs, const double* q, double* state8);
void orc_state2set(const orc_case* cs, int basis, const double* state8, double* q);
void orc_delta(const orc_case* cs, int from, int to, const double* state8, const double* dq, double* out);
void orc_roe_average(const orc_case* cs, const double* sl8, const double* sr8, double* out8);
int  orc_flux(const orc_case* cs, const orc_face* f, const double* ql, const double* qr, double* out);
double orc_limit(int limiter, double a, double b);
double orc_machsplit(int which, int sign, double m, double param);
void orc_bc_update(const orc_case* cs, int bctype, const orc_face* f, const double* cell3,
                   const double* qint, const double* qref, double* out);
void orc_wall_flux(const orc_case* cs, const orc_face* f, const double* qint, double* out);

/* least squares (lsq/xmetric.ipp, qmetric.ipp, factor_solve.ipp, spatial/lsqMetrics.h) */
void orc_xmetrics(const orc_case* cs, int ni, int nj, const double* cells, double* xm);
void orc_qmetrics(const orc_case* cs, int ni, int nj, const double* cells, const double* q, double* qm);
void orc_lsq_solve(const double* xm3, const double* qm, int nvar, const double* dir, double* out);
void orc_muscl_face(const orc_case* cs, const orc_face* f, const double* cl3, const double* cr3,
                    const double* ql, const double* qr, const double* xl, const double* xr,
                    const double* qml, const double* qmr, double* out);

/* field level (spatial/residualCalc.h, timestepping/rungeKutta.h) */
int orc_residual(const orc_case* cs, int ni, int nj, const double* nodes, const double* q,
                 const double* bref, double* res);
double orc_spectral_radius(int ni, int nj, const double* cells, const double* res, int nvar);
int orc_integrate(const orc_case* cs, int ni, int nj, const double* nodes, double* q,
                  const double* bref, int nsteps, double* t_out, double* timers_ms);
/* integrate with the accumulated stage residual smoothed (orc_smooth_residual) before each update */
int orc_integrate_smoothed(const orc_case* cs, int ni, int nj, const double* nodes, double* q,
                           const double* bref, int nsteps, int smooth_mode, double smooth_alpha, double* t_out);
void orc_norms(int ni, int nj, int nvar, const double* cells, const double* res,
               double* l2, double* linf);
void orc_set_threads(int n);
int  orc_get_threads(void);

/* legacy gradient (spatial/gradientCalc.h:16-313): dq [N][2][nvar] */
void orc_gradient(const orc_case* cs, int ni, int nj, const double* q, double* dq);
/* residual smoothing, product of the 1-D operators of OLD.fieldOperations/central*Smoothing.ipp;
   mode 0 explicit, 1 implicit, 2 explicit then implicit */
void orc_smooth_residual(const orc_case* cs, int ni, int nj, int mode, double alpha, double* res);

/* RK tables (ode.h) and the added classical RK4 */
void orc_rk_table(int which, int* nstages, double* alpha36, double* beta6);

#ifdef __cplusplus
}
#endif
#endif
