"""Case/driver layer: meshes, initial conditions, Runge-Kutta tables and case
descriptors of the benchmark configurations.

Mesh generators restate make_linspace_mesh / make_cylinder_mesh
(mesh/generate/twoD.h:15-40, 139-179) and add the geometrically stretched
O-grid of the cylinder cases.  Runge-Kutta tables are ODE::Explicit::ssp11/22/
33/34 (ode.h:107-197) plus classical RK4 in the same (alpha, beta) form.
"""
import math

import numpy as np

from . import (BASIS_CONSERVED, BASIS_PRIMITIVE, BC_PERIODIC, BC_RIEMANN, BC_WALL, DT_FIXED, DT_GLOBAL, DT_LOCAL,
               FLUX_AUSMPUP, FLUX_ROE, FLUX_RUSANOV, FLUX_SLAU, LAW_EULER, LAW_SCALAR, LIM_MC2, LIM_MINMOD2, LIM_NOLIMIT3,
               LIM_SMOOTHVANALBADA, LIM_VANALBADA2, LIM_VANLEER2, SCALE_ACOUSTIC, SCALE_CONVECTIVE, CaseDesc,
               Species)

GAMMA = 1.4


def air_species(minf=None):
    """get_air_species<double>() (conservationLaws/euler/species.ipp:2-18)."""
    s = Species()
    s.gamma = 1.4
    s.minf = 0.0001 if minf is None else minf
    s.lref = 1.0
    s.pr = 0.7
    s.nu = 1.81e-5
    s.dt = np.finfo(np.float64).max
    s.R = 287.058
    s.gamma1 = 1.0 / (s.gamma - 1.0)
    return s


def rk_table(name):
    """(nstages, alpha[6][6], beta[6]) of ssp11, ssp22, ssp33, ssp34 (ode.h:107-197) or rk4."""
    a = np.zeros((6, 6))
    b = np.zeros(6)
    if name == "ssp11":
        n = 1
        b[0] = 1.0
        a[0, 0] = 1.0
    elif name == "ssp22":
        n = 2
        b[:2] = (1.0, 1.0)
        a[0, 0] = 1.0
        a[1, :2] = (0.5, 0.5)
    elif name == "ssp33":
        n = 3
        b[:3] = (1.0, 0.5, 1.0)
        a[0, 0] = 1.0
        a[1, :2] = (0.5, 0.5)
        a[2, :3] = (1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0)
    elif name == "ssp34":
        n = 4
        b[:4] = (0.5, 1.0, 0.5, 1.0)
        a[0, 0] = 1.0
        a[1, :2] = (0.5, 0.5)
        a[2, :3] = (1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0)
        a[3, :4] = (1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 2.0)
    elif name == "rk4":
        # classical RK4: q_{s+1} = q0 + beta_s dt sum_k alpha[s][k] R_k
        n = 4
        b[:4] = (0.5, 0.5, 1.0, 1.0)
        a[0, 0] = 1.0
        a[1, 1] = 1.0
        a[2, 2] = 1.0
        a[3, :4] = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)
    else:
        raise ValueError(name)
    return n, a, b


def make_case(law, basis, flux, limiter, bc, rk, cfl, dt_mode=DT_GLOBAL, species=None, period_i=(0.0, 0.0),
              period_j=(0.0, 0.0), vscale=None, pscale=None):
    cs = CaseDesc()
    cs.law, cs.basis, cs.flux, cs.limiter = law, basis, flux, limiter
    cs.vscale = SCALE_CONVECTIVE if vscale is None else vscale
    cs.pscale = (SCALE_ACOUSTIC if flux == FLUX_SLAU else SCALE_CONVECTIVE) if pscale is None else pscale
    for k in range(4):
        cs.bc[k] = bc[k]
    n, a, b = rk_table(rk)
    cs.nstages = n
    for s in range(6):
        cs.beta[s] = b[s]
        for k in range(6):
            cs.alpha[s][k] = a[s, k]
    cs.cfl = cfl
    cs.dt_mode = dt_mode
    cs.period_i[0], cs.period_i[1] = period_i
    cs.period_j[0], cs.period_j[1] = period_j
    cs.sp = species if species is not None else air_species()
    return cs


# ---------------- meshes ----------------

def mesh_linspace(ni, nj, lox, hix, loy, hiy):
    """make_linspace_mesh (mesh/generate/twoD.h:15-40): nodes [ni+1][nj+1][2]."""
    dx = (hix - lox) / ni
    dy = (hiy - loy) / nj
    i = np.arange(ni + 1, dtype=np.float64)[:, None]
    j = np.arange(nj + 1, dtype=np.float64)[None, :]
    nodes = np.empty((ni + 1, nj + 1, 2))
    nodes[..., 0] = lox + i * dx
    nodes[..., 1] = loy + j * dy
    return nodes


def ogrid_ratio(nr, rc=0.5, rb=20.0, dr0=1e-3):
    """Geometric ratio s with dr0 (s^nr - 1)/(s - 1) = rb - rc (bisection)."""
    L = rb - rc

    def g(s):
        return dr0 * (s ** nr - 1.0) / (s - 1.0) - L

    lo, hi = 1.0 + 1e-12, 2.0
    while g(hi) < 0:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            break
        if g(mid) < 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def mesh_ogrid(nr, nth, rc=0.5, rb=20.0, dr0=1e-3):
    """O-grid around a cylinder, i = radius (geometric stretching, first spacing dr0),
    j = theta (uniform, node nth identical to node 0), as make_cylinder_mesh's layout
    (mesh/generate/twoD.h:139-179)."""
    s = ogrid_ratio(nr, rc, rb, dr0)
    i = np.arange(nr + 1)
    r = rc + dr0 * (np.power(s, i.astype(np.float64)) - 1.0) / (s - 1.0)
    r[nr] = rb
    jj = np.arange(nth + 1)
    jj[nth] = 0
    th = (2.0 * math.pi * jj) / nth
    nodes = np.empty((nr + 1, nth + 1, 2))
    nodes[..., 0] = r[:, None] * np.cos(th)[None, :]
    nodes[..., 1] = r[:, None] * np.sin(th)[None, :]
    return nodes


def cell_centres(nodes):
    """Vertex-average cell centres (geometry/volume.ipp:34-42)."""
    p0, p1, p2, p3 = nodes[:-1, :-1], nodes[1:, :-1], nodes[:-1, 1:], nodes[1:, 1:]
    return (((0.25 * p0 + 0.25 * p1) + 0.25 * p2) + 0.25 * p3)


def bref_size(ni, nj):
    return 2 * nj + 2 * ni


# ---------------- Euler helpers ----------------

def prim_to_cons(qp, gamma=GAMMA, R=287.058):
    """Conserved {rho u, rho v, rho, rho E} from primitive {u, v, rho, p}
    (euler/transforms/primitive.ipp:24-60 then conserved.ipp:2-29)."""
    qp = np.asarray(qp, dtype=np.float64)
    u, v, r, p = qp[..., 0], qp[..., 1], qp[..., 2], qp[..., 3]
    k = 0.5 * (u * u + v * v)
    t = p / (r * R)
    a2 = gamma * R * t
    h = a2 * (1.0 / (gamma - 1.0)) + k
    return np.stack([r * u, r * v, r, r * h - p], axis=-1)


def isentropic_vortex(x, y, t=0.0, beta=5.0, mach=0.5, gamma=GAMMA, x0=0.0, y0=0.0, L=20.0):
    """Shu's isentropic vortex (rho_inf = p_inf = 1) advected diagonally at Mach `mach`,
    primitive {u, v, rho, p}; periodic domain of side L centred on the origin."""
    a_inf = math.sqrt(gamma)
    uinf = mach * a_inf / math.sqrt(2.0)
    vinf = uinf
    xc = (x0 + uinf * t + L / 2) % L - L / 2
    yc = (y0 + vinf * t + L / 2) % L - L / 2
    dx = (x - xc + L / 2) % L - L / 2
    dy = (y - yc + L / 2) % L - L / 2
    r2 = dx * dx + dy * dy
    e = np.exp((1.0 - r2) / 2.0)
    du = -beta / (2.0 * math.pi) * e * dy
    dv = beta / (2.0 * math.pi) * e * dx
    T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * gamma * math.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return np.stack([uinf + du, vinf + dv, rho, p], axis=-1)


# ---------------- benchmark configurations ----------------

def config_scalar_advection(n=128):
    """Scalar advection on a periodic n x n grid over [0,1]^2, velocity (1, 0.5), Gaussian
    bump, Rusanov (= upwind for linear advection), VanLeer2, SSP-RK3, CFL 0.5, FP64."""
    nodes = mesh_linspace(n, n, 0.0, 1.0, 0.0, 1.0)
    c = cell_centres(nodes)
    q = np.empty((n, n, 3))
    q[..., 0] = 1.0
    q[..., 1] = 0.5
    q[..., 2] = np.exp(-((c[..., 0] - 0.5) ** 2 + (c[..., 1] - 0.5) ** 2) / 0.01)
    cs = make_case(LAW_SCALAR, BASIS_CONSERVED, FLUX_RUSANOV, LIM_VANLEER2, (BC_PERIODIC,) * 4, "ssp33", 0.5,
                   DT_GLOBAL, Species(), period_i=(1.0, 0.0), period_j=(0.0, 1.0))
    return dict(name="scalar-advection-128x128-vanleer2-upwind-ssprk3", nodes=nodes, q=q, case=cs,
                bref=np.zeros((bref_size(n, n), 3)), nsteps=200)


def config_cylinder_ogrid(nth=512, nr=256, mach=0.3, nsteps=2000):
    """Euler O-grid around a unit-diameter cylinder, 512 (theta) x 256 (r) cells, r from 0.5 to 20
    with first spacing 1e-3, freestream M=0.3 (rho=1, p=1/gamma), slip wall (id 0), Riemann far field
    (id 1), periodic theta (ids 2, 3); Roe + Harten fix, MinMod2, LSQ, RK4 with local dt at CFL 1."""
    nodes = mesh_ogrid(nr, nth)
    qinf = np.array([mach, 0.0, 1.0, 1.0 / GAMMA])
    q = np.broadcast_to(qinf, (nr, nth, 4)).copy()
    bref = np.broadcast_to(qinf, (bref_size(nr, nth), 4)).copy()
    cs = make_case(LAW_EULER, BASIS_PRIMITIVE, FLUX_ROE, LIM_MINMOD2, (BC_WALL, BC_RIEMANN, BC_PERIODIC, BC_PERIODIC),
                   "rk4", 1.0, DT_LOCAL, air_species(mach))
    return dict(name=f"euler-cylinder-ogrid-{nth}x{nr}-roe-minmod-lsq-rk4-local", nodes=nodes, q=q, case=cs,
                bref=bref, nsteps=nsteps)


def shocktube_states(problem="sod", species=None):
    """Left / right primitive (rho, u, p) of the shock tubes of cases/euler/oneD/shocktubes.h:
    Sod's (rho, u, p) = (1, 0, 1) | (0.125, 0, 0.1); the low-Mach tube (Sachdev, Hosangadi and
    Sankaran 2012) p = 100028.04 | 100000 Pa at T = 300 K, rho = p / (R T)."""
    if problem == "sod":
        return (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    R = (species or air_species()).R
    return (100028.04 / (R * 300.0), 0.0, 100028.04), (100000.0 / (R * 300.0), 0.0, 100000.0)


def config_shocktube(n, flux=FLUX_ROE, limiter=LIM_MC2, t_end=0.2, nsteps=None, problem="sod", rk="ssp34",
                     basis=BASIS_PRIMITIVE, minf=1.0):
    """The reference's 1-D shock tube (script/euler/shocktube.cpp: MUSCL + MC2, SSP(3,4), Riemann
    ghost cells, air species) on an n x 4 strip of square cells over [0, 1] x [0, 4/n], periodic
    across the strip, diaphragm at x = 1/2.  Fixed step so that the physical time is t_end: the
    update divides by the reference's volume (area / 4), so one step of dt advances 4 dt.
    minf (the species' free-stream Mach number, the low-Mach cut-off of AUSM+up and SLAU) is 1:
    there is no free stream, and the species default 1e-4 scales AUSM+up's pressure-diffusion
    mass flux by ~1/minf, which at the diaphragm blows the tube up within two steps."""
    sp = air_species(minf)
    left, right = shocktube_states(problem, sp)
    w = 4
    nodes = mesh_linspace(n, w, 0.0, 1.0, 0.0, w / n)
    x = cell_centres(nodes)[..., 0]
    q = np.empty((n, w, 4))
    for k, (r, u, p) in enumerate((left, right)):
        m = (x < 0.5) if k == 0 else (x >= 0.5)
        q[m] = (u, 0.0, r, p)
    if nsteps is None:
        nsteps = n
    bref = np.empty((bref_size(n, w), 4))
    bref[:] = (left[1], 0.0, left[0], left[2])          # boundary id 0 (x = 0); the strip's sides are periodic
    bref[w:2 * w] = (right[1], 0.0, right[0], right[2])   # boundary id 1 (x = 1)
    cs = make_case(LAW_EULER, basis, flux, limiter, (BC_RIEMANN, BC_RIEMANN, BC_PERIODIC, BC_PERIODIC), rk,
                   t_end / nsteps / 4.0, DT_FIXED, sp, period_j=(0.0, w / n))
    if basis == BASIS_CONSERVED:
        q = prim_to_cons(q)
        bref = prim_to_cons(bref)
    return dict(name=f"euler-shocktube-{problem}-{n}", nodes=nodes, q=q, case=cs, bref=bref, nsteps=nsteps,
                left=left, right=right, t_end=t_end, x=x[:, 0])


def gresho_state(x, y, mach=0.3, vel_inf=0.1, density=1.0, gamma=GAMMA):
    """Gresho vortex (script/euler/gresho2D.cpp:163-205) advected at vel_inf along x: azimuthal
    velocity 5r for r < 0.2, 2 - 5r for r < 0.4, 0 outside; pressure p0 + 12.5 r^2, then
    p0 + 12.5 r^2 + 4 (1 - 5r - ln 0.2 + ln r), then p0 - 2 + 4 ln 2, p0 = rho / (gamma M^2),
    so that dp/dr = rho u_theta^2 / r.  Primitive {u, v, rho, p}; r must not vanish."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    p0 = density / (gamma * mach * mach)
    r = np.sqrt(x * x + y * y)
    with np.errstate(divide="ignore", invalid="ignore"):
        lnr = np.log(r)
    inner, ring = r < 0.2, (r >= 0.2) & (r < 0.4)
    ut = np.where(inner, 5.0 * r, np.where(ring, 2.0 - 5.0 * r, 0.0))
    p = np.where(inner, p0 + 12.5 * r * r,
                 np.where(ring, p0 + 12.5 * r * r + 4.0 * (1.0 - 5.0 * r - math.log(0.2) + lnr),
                          p0 - 2.0 + 4.0 * math.log(2.0)))
    sin_t, cos_t = y / r, x / r
    q = np.empty(x.shape + (4,))
    q[..., 0] = vel_inf - ut * sin_t
    q[..., 1] = ut * cos_t
    q[..., 2] = density
    q[..., 3] = p
    return q


def config_gresho(n=256, flux=FLUX_RUSANOV, limiter=LIM_NOLIMIT3, nsteps=10, cfl=1.6, mach=0.3, vel_inf=0.1,
                  density=1.0):
    """The reference's Gresho driver (script/euler/gresho2D.cpp:78-110, 161-224): n x n cells over
    [-1, 1]^2, periodic, Rusanov + NoLimit3 MUSCL, SSP(3,4), global dt at CFL 1.6, species minf = M.
    qref is the outer (rest) state used for the boundary fill and the output's reference values."""
    nodes = mesh_linspace(n, n, -1.0, 1.0, -1.0, 1.0)
    c = cell_centres(nodes)
    q = gresho_state(c[..., 0], c[..., 1], mach, vel_inf, density)
    p0 = density / (GAMMA * mach * mach)
    qref = np.array([0.0, 0.0, density, p0 - 2.0 + 4.0 * math.log(2.0)])
    cs = make_case(LAW_EULER, BASIS_PRIMITIVE, flux, limiter, (BC_PERIODIC,) * 4, "ssp34", cfl, DT_GLOBAL,
                   air_species(mach), period_i=(2.0, 0.0), period_j=(0.0, 2.0))
    return dict(name=f"euler-gresho-{n}x{n}", nodes=nodes, q=q, case=cs, bref=np.broadcast_to(qref, (bref_size(n, n), 4)).copy(),
                qref=qref, nsteps=nsteps)


def config_soundwave(n, flux=FLUX_ROE, limiter=LIM_SMOOTHVANALBADA, mach=0.5, amplitude=0.01, nwavelengths=1,
                     cfl=0.5, crossings=1.0, rk="ssp34"):
    """The forward-travelling sound wave of cases/euler/oneD/periodic.h:37-80 on an n x 4
    periodic strip over [0, 1]: background u0 = rho0 = 1, a0 = u0 / M, p0 = rho0 a0^2 / gamma;
    p = p0 (1 + M A sin(k x)), rho from the isentrope, u from the u - 2a/(gamma-1) invariant of
    the u + a wave.  Fixed step of cfl h / (u0 + a0) (the update divides by the reference's
    area / 4 volume, so dt is a quarter of that), run for the given number of domain crossings."""
    sp = air_species(mach)
    w = 4
    nodes = mesh_linspace(n, w, 0.0, 1.0, 0.0, w / n)
    x = cell_centres(nodes)[..., 0]
    g = sp.gamma
    u0, r0 = 1.0, 1.0
    a0 = u0 / mach
    p0 = r0 * a0 * a0 / g
    k = 2.0 * math.pi * nwavelengths
    p = p0 * (1.0 + mach * amplitude * np.sin(k * x))
    r = r0 * np.power(p / p0, 1.0 / g)
    a = np.sqrt(g * p / r)
    q = np.empty((n, w, 4))
    q[..., 0] = u0 + 2.0 * (a - a0) / (g - 1.0)
    q[..., 1] = 0.0
    q[..., 2] = r
    q[..., 3] = p
    dt = cfl * (1.0 / n) / (u0 + a0)
    nsteps = int(round(crossings / (u0 + a0) / dt))
    cs = make_case(LAW_EULER, BASIS_PRIMITIVE, flux, limiter, (BC_PERIODIC,) * 4, rk, dt / 4.0, DT_FIXED, sp,
                   period_i=(1.0, 0.0), period_j=(0.0, w / n))
    return dict(name=f"euler-soundwave-{n}", nodes=nodes, q=q, case=cs, bref=np.zeros((bref_size(n, w), 4)),
                nsteps=nsteps, t_end=nsteps * dt)


def mesh_channelbump(d=2.0, l=3.0, w=2.0, h=0.1, nd=96, nl=48, nw=48):
    """make_channelbump_mesh (mesh/generate/twoD.h:49-130): a channel of height d with a cosine
    bump of height h and width w on its lower wall between two flat sections of length l;
    uniform cells in the flat sections, and over the bump nodes evenly spaced from the bump
    y = h/2 (1 + cos(2 pi x / w)) to the upper wall.  Nodes [2 nl + nw + 1][nd + 1][2]."""
    nx, ny = 2 * nl + nw, nd
    nodes = np.empty((nx + 1, ny + 1, 2))
    xl0, xw0, xr0 = -(l + 0.5 * w), -0.5 * w, 0.5 * w
    dxl, dyl = l / nl, d / ny
    for i in range(nl + 1):
        for j in range(ny + 1):
            nodes[i, j] = (xl0 + i * dxl, j * dyl)
            nodes[nl + nw + i, j] = (xr0 + i * dxl, j * dyl)
    dxw = w / nw
    for i in range(nw):
        ii = i + nl + 1
        x = xw0 + (i + 1) * dxw
        yl = 0.5 * h * (1.0 + math.cos(2.0 * math.pi * x / w))
        dpy = (d - yl) * (1.0 / ny)
        for j in range(ny):
            nodes[ii, j] = (x, yl + j * dpy)
        nodes[ii, ny] = (x, d)
    return nodes


def config_channelbump(nd=96, nl=48, nw=48, flux=FLUX_RUSANOV, limiter=None, nsteps=6500, cfl=0.4, mach=0.4,
                       rk="ssp11"):
    """The reference's channel-bump driver (script/euler/channelBump2D.cpp:80-140): d = 2, l = 3,
    w = 2, h = 0.1, free stream M = 0.4 (p = rho = 1) along x, fixed states at inflow and outflow
    (ids 0, 1), slip walls below and above (ids 2, 3), Rusanov + NoLimit1, forward Euler (ssp11),
    global dt at CFL 0.4, 6500 steps."""
    from . import BC_FIXED, LIM_NOLIMIT1
    nodes = mesh_channelbump(nd=nd, nl=nl, nw=nw)
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    speed = mach * math.sqrt(1.4 * 1.0 / 1.0)
    qref = np.array([speed, 0.0, 1.0, 1.0])
    q = np.broadcast_to(qref, (ni, nj, 4)).copy()
    cs = make_case(LAW_EULER, BASIS_PRIMITIVE, flux, LIM_NOLIMIT1 if limiter is None else limiter,
                   (BC_FIXED, BC_FIXED, BC_WALL, BC_WALL), rk, cfl, DT_GLOBAL, air_species(mach))
    return dict(name=f"euler-channelbump-{ni}x{nj}", nodes=nodes, q=q, case=cs,
                bref=np.broadcast_to(qref, (bref_size(ni, nj), 4)).copy(), qref=qref, nsteps=nsteps)


def config_isentropic_vortex(n=4096, flux=FLUX_SLAU, nsteps=500):
    """Isentropic vortex on a periodic n x n grid over [-10,10]^2, M=0.5 diagonal, strength 5;
    SLAU or AUSM+up, VanAlbada2, SSP-RK3, CFL 0.8."""
    nodes = mesh_linspace(n, n, -10.0, 10.0, -10.0, 10.0)
    c = cell_centres(nodes)
    q = isentropic_vortex(c[..., 0], c[..., 1])
    cs = make_case(LAW_EULER, BASIS_PRIMITIVE, flux, LIM_VANALBADA2, (BC_PERIODIC,) * 4, "ssp33", 0.8, DT_GLOBAL,
                   air_species(0.5), period_i=(20.0, 0.0), period_j=(0.0, 20.0))
    fname = "slau" if flux == FLUX_SLAU else "ausmpup"
    return dict(name=f"euler-isentropic-vortex-{n}x{n}-{fname}-vanalbada-ssprk3", nodes=nodes, q=q, case=cs,
                bref=np.zeros((bref_size(n, n), 4)), nsteps=nsteps)


# ---------------- batch of cylinder cases ----------------

class MT19937_64:
    """64-bit Mersenne Twister (Matsumoto & Nishimura 2004), the engine of std::mt19937_64:
    w=64, n=312, m=156, r=31, tempering (u,d,s,b,t,c,l) = (29, 0x5555555555555555, 17,
    0x71D67FFFEDA60000, 37, 0xFFF7EEE000000000, 43), seeding multiplier 6364136223846793005."""
    N, M = 312, 156
    MATRIX_A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF
    MASK = (1 << 64) - 1

    def __init__(self, seed=5489):
        mt = [0] * self.N
        mt[0] = seed & self.MASK
        for i in range(1, self.N):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & self.MASK
        self.mt, self.idx = mt, self.N

    def _twist(self):
        mt, N, M = self.mt, self.N, self.M
        for i in range(N):
            x = (mt[i] & self.UPPER) | (mt[(i + 1) % N] & self.LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self.MATRIX_A
            mt[i] = mt[(i + M) % N] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self.MASK


BATCH_SIZES = ((128, 64), (256, 128), (384, 192), (512, 256))   # (n_theta, n_r)


def batch_case_table(ncases=96, seed=20261202):
    """Cylinder batch: per case, in draw order, size index = rng() % 4, Mach = 0.2 + 0.6 U and
    angle of attack = -5 + 10 U degrees, U = (rng() >> 11) * 2^-53."""
    rng = MT19937_64(seed)
    rows = []
    for k in range(ncases):
        size = rng() % 4
        mach = 0.2 + 0.6 * ((rng() >> 11) * 2.0 ** -53)
        alpha = -5.0 + 10.0 * ((rng() >> 11) * 2.0 ** -53)
        nth, nr = BATCH_SIZES[size]
        rows.append(dict(case=k, size=size, nth=nth, nr=nr, mach=mach, alpha_deg=alpha))
    return rows


def config_cylinder(nth, nr, mach, alpha_deg=0.0, flux=FLUX_ROE, nsteps=1000):
    """One case of the cylinder batch: the O-grid of config_cylinder_ogrid at (n_theta, n_r), freestream
    (u, v) = M (cos a, sin a), rho = 1, p = 1/gamma; Roe, MinMod2, LSQ, RK4, local dt at CFL 1."""
    nodes = mesh_ogrid(nr, nth)
    a = math.radians(alpha_deg)
    qinf = np.array([mach * math.cos(a), mach * math.sin(a), 1.0, 1.0 / GAMMA])
    q = np.broadcast_to(qinf, (nr, nth, 4)).copy()
    bref = np.broadcast_to(qinf, (bref_size(nr, nth), 4)).copy()
    cs = make_case(LAW_EULER, BASIS_PRIMITIVE, flux, LIM_MINMOD2, (BC_WALL, BC_RIEMANN, BC_PERIODIC, BC_PERIODIC),
                   "rk4", 1.0, DT_LOCAL, air_species(mach))
    return dict(name=f"euler-cylinder-{nth}x{nr}-M{mach:.3f}-a{alpha_deg:+.2f}", nodes=nodes, q=q, case=cs,
                bref=bref, nsteps=nsteps)


# ---------------- adjoint case: cylinder, AUSM+up, smooth van Albada ----------------

def smooth_perturbation(nr, nth, seed=4242, amp=1e-3, modes=4):
    """Smooth field on the O-grid, per conserved variable k:
    amp * sum_{m,n=1..modes} a_kmn sin(pi n (i+1/2)/nr) cos(m theta_j + phi_kmn), theta_j = 2 pi (j+1/2)/nth,
    a ~ U[-1,1]/16 and phi ~ U[0, 2 pi) drawn from MT19937-64(seed) in the order k, m, n, (a, phi)."""
    rng = MT19937_64(seed)
    u = lambda: (rng() >> 11) * 2.0 ** -53
    i = (np.arange(nr) + 0.5)[:, None]
    th = (2.0 * math.pi * (np.arange(nth) + 0.5) / nth)[None, :]
    out = np.zeros((nr, nth, 4))
    for k in range(4):
        for m in range(1, modes + 1):
            for n in range(1, modes + 1):
                a = (2.0 * u() - 1.0) / 16.0
                phi = 2.0 * math.pi * u()
                out[..., k] += a * np.sin(math.pi * n * i / nr) * np.cos(m * th + phi)
    return amp * out


def freestream_cons(mach, alpha_deg=0.0):
    a = math.radians(alpha_deg)
    return prim_to_cons(np.array([mach * math.cos(a), mach * math.sin(a), 1.0, 1.0 / GAMMA]))


def config_adjoint(nth=1024, nr=512, mach=0.4, nsteps=400, checkpoint_every=20, dt=None):
    """Cylinder O-grid (the C2 rule at n_theta x n_r), M = 0.4, AUSM+up, smooth van Albada,
    SSP-RK3 with a fixed time step, conserved basis; initial state = free stream + the seeded
    smooth perturbation.  dt = None leaves the step to the caller (CFL 0.8 of the initial state)."""
    nodes = mesh_ogrid(nr, nth)
    qinf = freestream_cons(mach)
    q = qinf[None, None, :] + smooth_perturbation(nr, nth)
    bref = np.broadcast_to(qinf, (bref_size(nr, nth), 4)).copy()
    cs = make_case(LAW_EULER, BASIS_CONSERVED, FLUX_AUSMPUP, LIM_SMOOTHVANALBADA,
                   (BC_WALL, BC_RIEMANN, BC_PERIODIC, BC_PERIODIC), "ssp33", 1.0 if dt is None else dt, DT_FIXED,
                   air_species(mach))
    return dict(name=f"euler-cylinder-{nth}x{nr}-ausmpup-smoothvanalbada-ssprk3-adjoint", nodes=nodes, q=q, case=cs,
                bref=bref, nsteps=nsteps, checkpoint_every=checkpoint_every, mach=mach)


def cell_volumes(nodes):
    """Cell volumes as the reference defines them: |J(0,0)| of the bilinear map = area / 4
    (geometry/volume.ipp); shoelace area of the cell quadrilateral."""
    p0, p1, p2, p3 = nodes[:-1, :-1], nodes[1:, :-1], nodes[1:, 1:], nodes[:-1, 1:]
    x = [p0[..., 0], p1[..., 0], p2[..., 0], p3[..., 0]]
    y = [p0[..., 1], p1[..., 1], p2[..., 1], p3[..., 1]]
    area = 0.5 * np.abs(sum(x[k] * y[(k + 1) % 4] - x[(k + 1) % 4] * y[k] for k in range(4)))
    return area / 4.0


def wall_faces(nodes):
    """Area and unit normal (into the fluid) of the wall faces on node line 0
    (geometry/surface.ipp:12-29 with the boundary orientation of boundaryFluxLoop.h:162-175)."""
    p0, p1 = nodes[0, :-1], nodes[0, 1:]
    t0 = p1 - p0
    l = np.sqrt(t0[:, 0] * t0[:, 0] + t0[:, 1] * t0[:, 1])
    s1 = (p0 - p1) * (1.0 / l)[:, None]
    return l, -s1[:, 1], s1[:, 0]


def wall_force_x(nodes, q, gamma=GAMMA):
    """J = -sum over wall faces of p_w A n_x (n into the fluid), p_w the pressure of the wall cell
    (conserved basis, inviscidWall.ipp: the wall state is the interior state)."""
    area, nx, _ = wall_faces(nodes)
    qw = q[0]
    p = (gamma - 1.0) * (qw[:, 3] - 0.5 * (qw[:, 0] ** 2 + qw[:, 1] ** 2) / qw[:, 2])
    return -np.sum(p * area * nx)


def wall_force_x_grad(nodes, q, gamma=GAMMA):
    """dJ/dq of wall_force_x (nonzero on the wall cells only)."""
    area, nx, _ = wall_faces(nodes)
    qw = q[0]
    u, v = qw[:, 0] / qw[:, 2], qw[:, 1] / qw[:, 2]
    g = np.zeros_like(q)
    w = -(area * nx) * (gamma - 1.0)
    g[0, :, 0] = w * (-u)
    g[0, :, 1] = w * (-v)
    g[0, :, 2] = w * (0.5 * (u * u + v * v))
    g[0, :, 3] = w
    return g

