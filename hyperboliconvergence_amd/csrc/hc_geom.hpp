// hc_geom.hpp -- one-time geometry of a structured mesh for the stage kernel:
// cell volumes and centres (geometry/volume.ipp:25-66, mesh/dual.ipp:42-83),
// face surfaces (geometry/surface.ipp:12-29), the LSQ spatial metric
// (spatial/lsqMetrics.h:31-163) and, per face, everything of the MUSCL
// reconstruction that depends on geometry only (spatial/muscl.h:35-50,
// lsq/factor_solve.ipp:13-100): distances d_L, d_R, 1/(d_L+d_R), the centre
// difference dx_c and the factored left/right biased metrics.
#pragma once

#include <vector>

namespace hc {

struct GeomInput {
    int ni, nj;
    const double* nodes;   // [ni+1][nj+1][2]
    int bc[4];             // boundary ids 0..3 (0 = periodic)
    double period_i[2];    // translation of a cell to its periodic image across boundary 1
    double period_j[2];    // ... across boundary 3
};

struct GeomOutput {
    std::vector<double> cvol, ccx, ccy;   // [ni*nj]
    std::vector<double> X;                // [ni*nj][3]  (kept for inspection)
    std::vector<double> fi;               // [NSLOT][(ni+1)*nj]
    std::vector<double> fj;               // [NSLOT][ni*(nj+1)]
};

// returns 0 on success, -1 on a degenerate cell or face
int build_geometry(const GeomInput& in, GeomOutput& out);

// Uniform geometry: the mesh is periodic in both directions and every face slot of the faces the
// stage kernel reads (i-faces at node lines 0 .. ni-1, j-faces at node lines 0 .. nj-1) is bitwise
// equal to that of face 0 of its kind, and every cell volume to that of cell 0 (a uniform Cartesian
// grid of a periodic box).  The stage kernel then takes the slots of face 0 and the volume of cell 0
// for all faces and cells, with the same bits as the per-face arrays.
bool uniform_geometry(const GeomInput& in, const GeomOutput& g);

// FP32 face slots of the mixed path from the FP64 ones: every slot rounded, except that the two
// biased LSQ factors become their directional weights w = X'^-1 n (two per side: the slope along
// n is Q'.w), formed in FP64 and rounded
void fp32_slots(const std::vector<double>& f64, long long nf, std::vector<float>& f32);

}   // namespace hc
