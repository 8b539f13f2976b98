// hc_tiles.hpp -- tile shapes and work-group sizes of the stage kernels (shared by the
// launchers, the context's autotuner and the host emulation of the kernels).
#pragma once

namespace hc {

struct TileShape {
    int ti, tj, nt;   // cells along i and j, lanes of the work-group
};

// Two shapes per precision; a context times both on its own cases and keeps the faster.
// The face phase runs 2*ti*tj + ti + tj faces on nt lanes:
//   FP64 4x32 on 256 lanes: 292 faces (one wave takes two), 128 cells per work-group; at
//        4 work-groups per CU a 131072-cell grid is exactly one round of work-groups
//   FP64 7x16 on 256 lanes: 247 faces, one per lane, every wave equally loaded
//   mixed 8x32 on 256 lanes: 552 faces; mixed 7x32 on 256 lanes: 487 faces, two per lane
constexpr int NTILECFG = 2;
constexpr TileShape TILES64[NTILECFG] = {{4, 32, 256}, {7, 16, 256}};
constexpr TileShape TILES32[NTILECFG] = {{8, 32, 256}, {7, 32, 256}};

}   // namespace hc
