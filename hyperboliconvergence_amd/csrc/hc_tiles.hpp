// hc_tiles.hpp -- tile shapes and work-group sizes of the stage kernels (shared by the
// launchers, the context's autotuner and the host emulation of the kernels).
#pragma once

#include "hc_point.hpp"

namespace hc {

struct TileShape {
    int ti, tj, nt;   // cells along i and j, lanes of the work-group
    int cst;          // centre staging: centre differences formed from cell centres in LDS
    int gpf;          // geometry prefetch: each lane's first face slots loaded before the metric phase
};

// Three configurations per precision; a context times each on its own cases and keeps the
// fastest (results are bit-identical).  The face phase runs 2*ti*tj + ti + tj faces on nt lanes:
//   4x32 on 256 lanes: 292 faces (one wave takes two), 128 cells per work-group; at 4
//        work-groups per CU a 131072-cell grid is exactly one round of work-groups
//   7x16 on 256 lanes: 247 faces, one per lane, every wave equally loaded
// Centre staging reads the cell centres of the q tile (16 B per halo cell) instead of two face
// slots per metric edge and face (about 10 % fewer bytes and load instructions): it pays on
// large grids, where work-groups of many rounds overlap their phases, and costs on a grid of
// one round, where every work-group's load phase is on the critical path.  Geometry prefetch
// issues each lane's first-face slot loads together with the tile load, so their latency is not
// paid again at the start of the face phase (the 7x16 shapes, one face per lane; on 4x32 the
// held slots would cost a wave of occupancy).
// The mixed path of the Roe / Rusanov / central fluxes keeps its solution tile and face fluxes
// in FP64 (mixed_wide), so it has the FP64 LDS budget and shapes; the all-FP32 faces of SLAU and
// AUSM+up take larger tiles: 8x32 (552 faces), 7x32 (487 faces, two per lane) on 256 lanes, or
// 7x16.
constexpr int NTILECFG = 3;
constexpr TileShape TILES64[NTILECFG] = {{4, 32, 256, 0, 0}, {7, 16, 256, 0, 1}, {7, 16, 256, 1, 1}};
constexpr TileShape TILES32[NTILECFG] = {{8, 32, 256, 0, 0}, {7, 32, 256, 0, 0}, {7, 16, 256, 0, 0}};

constexpr const TileShape* tile_shapes(bool mixed, int law, int flux)
{
    return (!mixed || mixed_wide(law, flux)) ? TILES64 : TILES32;
}

}   // namespace hc
