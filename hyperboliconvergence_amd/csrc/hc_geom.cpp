// hc_geom.cpp -- one-time host geometry of a structured mesh (see hc_geom.hpp).
// Built with -ffp-contract=off: every operation is one correctly rounded IEEE
// operation in the reference's evaluation order.
#include "hc_geom.hpp"
#include "hc_stage.hpp"

#include <cmath>
#include <cstring>

namespace hc {
namespace {

struct Surf {
    double area, cx, cy, n0, n1;
};

// geometry/surface.ipp:12-29 (metric[0] = orthog(tangent1 / l))
Surf make_surface(const double* p0, const double* p1)
{
    const double t0x = p1[0] - p0[0], t0y = p1[1] - p0[1];
    const double t1x = p0[0] - p1[0], t1y = p0[1] - p1[1];
    double l2 = 0;
    l2 = l2 + t0x * t0x;
    l2 = l2 + t0y * t0y;
    const double l = std::sqrt(l2);
    const double l1 = 1. / l;
    Surf s;
    s.area = l;
    s.cx = p0[0] + t0x * 0.5;
    s.cy = p0[1] + t0y * 0.5;
    const double a = t1x * l1, b = t1y * l1;
    s.n0 = -b;
    s.n1 = a;
    return s;
}

inline double length(double dx, double dy)
{
    double d = 0;
    d = d + dx * dx;
    d = d + dy * dy;
    return std::sqrt(d);
}

}   // namespace

int build_geometry(const GeomInput& in, GeomOutput& out)
{
    const int ni = in.ni, nj = in.nj;
    const long long N = (long long)ni * nj;
    const long long nfi = (long long)(ni + 1) * nj;
    const long long nfj = (long long)ni * (nj + 1);
    auto node = [&](int i, int j) { return in.nodes + 2 * ((long long)i * (nj + 1) + j); };
    out.cvol.assign(N, 0.);
    out.ccx.assign(N, 0.);
    out.ccy.assign(N, 0.);
    out.X.assign(3 * N, 0.);
    out.fi.assign(NSLOT * nfi, 0.);
    out.fj.assign(NSLOT * nfj, 0.);
    const bool per_i = in.bc[0] == BC_PERIODIC;
    const bool per_j = in.bc[2] == BC_PERIODIC;

    // cells: volume.ipp:25-66 evaluated at the element centre (shape values 1/4,
    // derivative rows (-1/4,-1/4), (1/4,-1/4), (-1/4,1/4), (1/4,1/4))
    int bad = 0;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
            const double* p[4] = {node(i, j), node(i + 1, j), node(i, j + 1), node(i + 1, j + 1)};
            const double u = 0.5 * 0.5;
            const double du[4][2] = {{-0.5 * 0.5, 0.5 * -0.5}, {0.5 * 0.5, 0.5 * -0.5}, {-0.5 * 0.5, 0.5 * 0.5}, {0.5 * 0.5, 0.5 * 0.5}};
            double com[2];
            for (int d = 0; d < 2; ++d) {
                double c = 0.;
                for (int k = 0; k < 4; ++k) c = c + u * p[k][d];
                com[d] = c;
            }
            double t[2][2];
            for (int a = 0; a < 2; ++a)
                for (int d = 0; d < 2; ++d) {
                    double c = 0.;
                    for (int k = 0; k < 4; ++k) c = c + du[k][a] * p[k][d];
                    t[a][d] = c;
                }
            const double vol = std::fabs(t[0][0] * t[1][1] - t[0][1] * t[1][0]);
            if (!(vol > 0.) || !std::isfinite(vol)) bad = 1;
            const long long c = (long long)i * nj + j;
            out.cvol[c] = vol;
            out.ccx[c] = com[0];
            out.ccy[c] = com[1];
        }
    if (bad) return -1;

    auto cx = [&](int i, int j) { return out.ccx[(long long)i * nj + j]; };
    auto cy = [&](int i, int j) { return out.ccy[(long long)i * nj + j]; };
    auto addX = [&](int i, int j, double dx, double dy) {
        double* x = &out.X[3 * ((long long)i * nj + j)];
        x[0] = x[0] + dx * dx;
        x[1] = x[1] + dx * dy;
        x[2] = x[2] + dy * dy;
    };
    // spatial metric, per cell: interior i-edges, interior j-edges, then the boundary
    // sides (periodic image edge, or the mirror point of lsqMetrics.h:116-124)
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni - 1; ++i) {
            const double dx = cx(i + 1, j) - cx(i, j), dy = cy(i + 1, j) - cy(i, j);
            addX(i, j, dx, dy);
            addX(i + 1, j, dx, dy);
        }
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj - 1; ++j) {
            const double dx = cx(i, j + 1) - cx(i, j), dy = cy(i, j + 1) - cy(i, j);
            addX(i, j, dx, dy);
            addX(i, j + 1, dx, dy);
        }
    auto mirror = [&](int i, int j, int i1, int j1) {
        const double c0x = cx(i, j), c0y = cy(i, j);
        const double dpx = c0x - cx(i1, j1), dpy = c0y - cy(i1, j1);
        const double bx = c0x + dpx, by = c0y + dpy;
        addX(i, j, bx - c0x, by - c0y);
    };
    for (int j = 0; j < nj; ++j) {
        if (per_i) {
            const double lx = cx(ni - 1, j) - in.period_i[0], ly = cy(ni - 1, j) - in.period_i[1];
            const double dx = cx(0, j) - lx, dy = cy(0, j) - ly;
            addX(ni - 1, j, dx, dy);
            addX(0, j, dx, dy);
        } else {
            mirror(0, j, 1, j);
            mirror(ni - 1, j, ni - 2, j);
        }
    }
    for (int i = 0; i < ni; ++i) {
        if (per_j) {
            const double lx = cx(i, nj - 1) - in.period_j[0], ly = cy(i, nj - 1) - in.period_j[1];
            const double dx = cx(i, 0) - lx, dy = cy(i, 0) - ly;
            addX(i, nj - 1, dx, dy);
            addX(i, 0, dx, dy);
        } else {
            mirror(i, 0, i, 1);
            mirror(i, nj - 1, i, nj - 2);
        }
    }

    // faces
    auto muscl_face = [&](double* base, long long nf, long long idx, const Surf& s, double clx, double cly, int il, int jl,
                          int ir, int jr) {
        const double crx = cx(ir, jr), cry = cy(ir, jr);
        base[FS_AREA * nf + idx] = s.area;
        base[FS_N0 * nf + idx] = s.n0;
        base[FS_N1 * nf + idx] = s.n1;
        const double dxc0 = crx - clx, dxc1 = cry - cly;
        const double dl = length(s.cx - clx, s.cy - cly);
        const double dr = length(crx - s.cx, cry - s.cy);
        base[FS_DL * nf + idx] = dl;
        base[FS_DR * nf + idx] = dr;
        base[FS_S1 * nf + idx] = 1. / (dl + dr);
        base[FS_DX0 * nf + idx] = dxc0;
        base[FS_DX1 * nf + idx] = dxc1;
        const double m0 = dxc0 * dxc0, m1 = dxc0 * dxc1, m2 = dxc1 * dxc1;
        const double* XL = &out.X[3 * ((long long)il * nj + jl)];
        const double* XR = &out.X[3 * ((long long)ir * nj + jr)];
        const double bl[3] = {XL[0] - m0, XL[1] - m1, XL[2] - m2};
        const double br[3] = {XR[0] - m0, XR[1] - m1, XR[2] - m2};
        double fl[3], fr[3];
        lsq_factor(bl, fl);
        lsq_factor(br, fr);
        base[FS_FL0 * nf + idx] = fl[0];
        base[FS_FL1 * nf + idx] = fl[1];
        base[FS_FL2 * nf + idx] = fl[2];
        base[FS_FR0 * nf + idx] = fr[0];
        base[FS_FR1 * nf + idx] = fr[1];
        base[FS_FR2 * nf + idx] = fr[2];
    };
    auto boundary_face = [&](double* base, long long nf, long long idx, const Surf& s) {
        base[FS_AREA * nf + idx] = s.area;
        base[FS_N0 * nf + idx] = s.n0;
        base[FS_N1 * nf + idx] = s.n1;
        base[FS_CX * nf + idx] = s.cx;
        base[FS_CY * nf + idx] = s.cy;
    };
    double* fi = out.fi.data();
    double* fj = out.fj.data();
    // i-faces: node line l, surface(nodes(l, j), nodes(l, j+1)) (residualCalc.h:211-214)
    for (int l = 0; l <= ni; ++l)
        for (int j = 0; j < nj; ++j) {
            const long long idx = (long long)l * nj + j;
            const Surf s = make_surface(node(l, j), node(l, j + 1));
            if (l >= 1 && l <= ni - 1) {
                muscl_face(fi, nfi, idx, s, cx(l - 1, j), cy(l - 1, j), l - 1, j, l, j);
            } else if (per_i) {
                if (l == 0) {
                    // periodicLoop.h:129-152: wrap face on node line 0, left cell is the image of ni-1
                    const double lx = cx(ni - 1, j) - in.period_i[0], ly = cy(ni - 1, j) - in.period_i[1];
                    muscl_face(fi, nfi, idx, s, lx, ly, ni - 1, j, 0, j);
                }
            } else {
                boundary_face(fi, nfi, idx, s);
            }
        }
    // j-faces: node line l, surface(nodes(i+1, l), nodes(i, l)) (residualCalc.h:239-242)
    for (int i = 0; i < ni; ++i)
        for (int l = 0; l <= nj; ++l) {
            const long long idx = (long long)i * (nj + 1) + l;
            const Surf s = make_surface(node(i + 1, l), node(i, l));
            if (l >= 1 && l <= nj - 1) {
                muscl_face(fj, nfj, idx, s, cx(i, l - 1), cy(i, l - 1), i, l - 1, i, l);
            } else if (per_j) {
                if (l == 0) {
                    const double lx = cx(i, nj - 1) - in.period_j[0], ly = cy(i, nj - 1) - in.period_j[1];
                    muscl_face(fj, nfj, idx, s, lx, ly, i, nj - 1, i, 0);
                }
            } else {
                boundary_face(fj, nfj, idx, s);
            }
        }
    return 0;
}

bool uniform_geometry(const GeomInput& in, const GeomOutput& g)
{
    const int ni = in.ni, nj = in.nj;
    if (in.bc[0] != BC_PERIODIC || in.bc[2] != BC_PERIODIC || ni < 1 || nj < 1) return false;
    auto same = [](double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; };
    const long long N = (long long)ni * nj;
    for (long long c = 1; c < N; ++c)
        if (!same(g.cvol[c], g.cvol[0])) return false;
    const long long nfi = (long long)(ni + 1) * nj, nfj = (long long)ni * (nj + 1);
    for (int s = 0; s < NSLOT; ++s) {
        const double* fi = g.fi.data() + s * nfi;
        const double* fj = g.fj.data() + s * nfj;
        for (int l = 0; l < ni; ++l)
            for (int t = 0; t < nj; ++t)
                if (!same(fi[(long long)l * nj + t], fi[0])) return false;
        for (int i = 0; i < ni; ++i)
            for (int l = 0; l < nj; ++l)
                if (!same(fj[(long long)i * (nj + 1) + l], fj[0])) return false;
    }
    return true;
}

void fp32_slots(const std::vector<double>& f64, long long nf, std::vector<float>& f32)
{
    f32.assign(f64.begin(), f64.end());
    for (long long k = 0; k < nf; ++k) {
        const double n0 = f64[FS_N0 * nf + k], n1 = f64[FS_N1 * nf + k];
        for (int side = 0; side < 2; ++side) {
            const int s0 = side == 0 ? FS_FL0 : FS_FR0;
            const double x0 = f64[s0 * nf + k], x1 = f64[(s0 + 1) * nf + k], x2 = f64[(s0 + 2) * nf + k];
            f32[s0 * nf + k] = float(x0 * n0 + x1 * n1);
            f32[(s0 + 1) * nf + k] = float(x1 * n0 + x2 * n1);
            f32[(s0 + 2) * nf + k] = 0.0f;
        }
    }
}

}   // namespace hc
