// ob_tiled3d.hpp -- tile-sorted 3D deposition (see ob_tiled3d.hip).
#pragma once

#include "ob_internal.hpp"

namespace ob {

// Grid of the deposit plus its xy tiling: tiles of T x T centre nodes, all nz z cells.
struct TileGeom {
  int nx, ny, nz;
  int T, tx, ty, ntiles;
  double xMin, xMax, yMin, yMax, zMin, zMax, dx, dy, dz;
};

// Picks the largest tile whose (T+2)^2 * nz LDS block fits lds_budget bytes (T = 0 if none).
TileGeom make_tile_geom(int nx, int ny, int nz, double xMin, double xMax, double yMin,
                        double yMax, double zMin, double zMax, double dx, double dy, double dz,
                        int lds_budget);
size_t tile_lds_bytes(const TileGeom& g);

// Scratch of the sort (owned by a calculator, grown on demand).
struct TileWork {
  uint32_t* key = nullptr;
  uint32_t *counts = nullptr, *offsets = nullptr, *cursor = nullptr;
  int4* items = nullptr;
  uint32_t* nitems = nullptr;
  int64_t n_cap = 0, t_cap = 0, i_cap = 0;
  void reserve(int64_t n, int ntiles);
  void release();
};

// Sort the bunch by tile (device order only) and deposit its charge into rho, a
// z-fastest [nx][ny][nz] array that must already be zero.
void sort_and_deposit_3d(ob_bunch* b, const TileGeom& g, double macrosize, double* rho,
                         TileWork& w);

}  // namespace ob
