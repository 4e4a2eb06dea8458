// ob_tiled3d.hpp -- tile-sorted 3D deposition and kick (see ob_tiled3d.hip).
#pragma once

#include "ob_internal.hpp"

namespace ob {

// Grid of the deposit plus its tiling into boxes of T x T x Tz centre nodes.  Each tile owns
// an LDS block of (T+2) x (T+2) x (Tz + 2*zh) nodes: the box plus the one-node halo the
// 3-point stencils reach (zh = 1 for nz >= 3; for nz <= 2 one z tile spans the grid).
struct TileGeom {
  int nx, ny, nz;
  int T, Tz, zh, tx, ty, tz, ntiles;
  double xMin, xMax, yMin, yMax, zMin, zMax, dx, dy, dz;
};

// Picks a tile shape whose LDS block fits lds_budget bytes and whose tile count fits the
// LDS histograms of the sort (T = 0 if none does).
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
// z-fastest [nx][ny][nz] array that must already be zero.  The work items stay valid for
// tile_kick_3d until the particles move.
void sort_and_deposit_3d(ob_bunch* b, const TileGeom& g, double macrosize, double* rho,
                         TileWork& w);

// The SpaceChargeCalc3D kick (SpaceChargeCalc3D.cc:120-147) of a bunch sorted by
// sort_and_deposit_3d with the same geometry: phi is the rest-frame potential on the same
// [nx][ny][nz] nodes, z grid (rzMin, rdz), bunch centre zc.
struct Kick3D {
  const double* phi;
  double rzMin, rdz, zc, gamma, trans_factor, long_factor;
  // SpaceChargeCalcSliceBySlice2D with a boundary: particles outside it are not kicked
  // (SpaceChargeCalcSliceBySlice2D.cc:104); type 0 = no boundary (SC2DState::bnd_type).
  int bnd_type;
  double bnd[6];
};
void tile_kick_3d(ob_bunch* b, const TileGeom& g, const Kick3D& k, TileWork& w);

}  // namespace ob
