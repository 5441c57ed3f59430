// Locally refined (quadtree) meshes over the terrain-following base mesh
// (host side, amr.cpp). The reference builds its mesh with p4est and
// refine_global (Atmospheric_Flow.cc:279-289); this is the build-defined
// local refinement: every base cell (bi, bj) of the nx0 x nz0 mesh is the
// root of a quadtree, a cell is (bi, bj, level l, sx, sz) with
// 0 <= sx, sz < 2^l, and its vertices are the images of the reference
// sub-box [sx, sx+1]/2^l x [sz, sz+1]/2^l under the base cell's bilinear
// map (the children deal.II creates for straight-edged cells).
//
// Leaves are given base cell by base cell in the base mesh's order
// (bj*nx0 + bi); their order is the cell order of every field. Neighbouring
// cells may differ by at most one level (2:1 balance across faces).
//
// The multigrid hierarchy coarsens globally: on each coarser level every
// cell whose three siblings are also cells of the level is replaced by its
// parent (which takes the place of its first child in the order), until
// the base mesh remains; below it the base mesh coarsens uniformly.
#pragma once
#include <vector>

#include "engine.h"

namespace atmo {

struct AmrLevelHost {
  int ncell = 0;
  bool base = false;           // the base mesh itself (structured, base-cell order)
  std::vector<QCell> cells;
  std::vector<double> cgeo;    // 6 per cell: x0, dx, z00, z10, z01, z11
  std::vector<int> cface;      // 16 per cell (finest level only), see MeshDev
  std::vector<int> child;      // 4 per cell: children in the next finer level (empty on the finest)
  std::vector<int> parent;     // 2 per cell: (cell, a + 2b or -1) in the next coarser level (empty on the base)
};

// Levels of a locally refined mesh, finest first, ending with the base mesh.
// Throws std::invalid_argument for cells out of range, duplicates, a
// non-covering or unbalanced leaf set, or leaves out of base-cell order.
std::vector<AmrLevelHost> build_amr_levels(const MeshDesc& base, const std::vector<QCell>& leaves);

// Longest diagonal minimum over the cells (GridTools::minimal_cell_diameter).
double amr_min_cell_diameter(const AmrLevelHost& finest);

}  // namespace atmo
