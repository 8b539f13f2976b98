// TEST-ONLY: evaluates the drop-in API's host-side point functors (hcamd.hpp) on the states
// read from stdin and prints the FluxResult, for comparison with the oracle.
//   input lines: flux_id limiter_id nx ny area uL vL rL pL uR vR rR pR
#include <hcamd/hcamd.hpp>

#include <cstdio>

constexpr LawType Law = LawType::Euler;
using Real = double;
using Prim = VariableSet<Law, 2, EulerBases::Primitive, Real>;

template <typename F> void eval(const F& f, const Species<Law, Real>& sp, const geom::Surface<2, Real>& s, const Prim& l, const Prim& r)
{
    const auto fr = f(sp, s, l, r);
    std::printf("%.17e %.17e %.17e %.17e %.17e\n", fr.flux[0], fr.flux[1], fr.flux[2], fr.flux[3], fr.lambda);
}

int main()
{
    Species<Law, Real> sp = get_air_species<Real>();
    sp.minf = 0.3;
    int fid, lid;
    double nx, ny, area;
    Prim l, r;
    while (std::scanf("%d %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &fid, &lid, &nx, &ny, &area, &l[0], &l[1], &l[2],
                      &l[3], &r[0], &r[1], &r[2], &r[3]) == 13) {
        geom::Surface<2, Real> s{};
        s.area = area;
        s.metric[0][0] = nx;
        s.metric[0][1] = ny;
        s.metric[1][0] = -ny;
        s.metric[1][1] = nx;
        switch (fid) {
        case 0: eval(CentralFlux<Law>{}, sp, s, l, r); break;
        case 1: eval(RusanovFlux<Law>{}, sp, s, l, r); break;
        case 2: eval(RoeFlux<Law>{}, sp, s, l, r); break;
        case 3: eval(Slau<>{}, sp, s, l, r); break;
        case 5: eval(RoeFluxFaceFrame<Law>{}, sp, s, l, r); break;
        case 6: eval(RoeUnprecWS{}, sp, s, l, r); break;
        default: eval(AusmPlusUP<>{}, sp, s, l, r); break;
        }
        // the limiter functor on the first two components
        using Del = VariableDelta<Law, 2, EulerBases::Primitive, Real>;
        Del a{{l[0], l[1], l[2], l[3]}}, b{{r[0], r[1], r[2], r[3]}};
        Del z;
        switch (lid) {
        case 4: z = Limiters::MinMod2{}(a, b); break;
        case 3: z = Limiters::VanAlbada2{}(a, b); break;
        default: z = Limiters::VanLeer2{}(a, b); break;
        }
        std::printf("%.17e %.17e %.17e %.17e\n", z[0], z[1], z[2], z[3]);
        // set2State of the left set with the reference's accessors
        const auto st = set2State(sp, l);
        std::printf("%.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e\n", st.pressure(), st.density(), st.temperature(),
                    st.specificTotalEnthalpy(), st.velocity2(), st.speedOfSound2(), st.velocity(0), st.velocity(1));
    }
    return 0;
}
