// ref_harness.cc -- C entry points over the reference's own Grid1D/Grid2D/Grid3D and
// Base/ShapedBoundary2D classes, compiled from the reference sources where they lie (see
// oracle/Makefile.ref).  TEST INFRASTRUCTURE ONLY: it writes golden vectors and
// checks the oracle's restatement of the same functions bit for bit.
//
// Only the Grid classes build here: Bunch (and so teapotbase and
// BunchExtremaCalculator) needs the Python-2 C-API MPI communicator wrapper, and
// PoissonSolverFFT2D/3D need fftw3.h; neither exists in this image.
#include <cmath>
#include "Grid1D.hh"
#include "Grid2D.hh"
#include "Grid3D.hh"
#include "ShapedBoundary2D.hh"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {
// Grid3D::getGridIndAndFrac is protected; a derived class can call it.
struct Grid3DProbe : public Grid3D {
  Grid3DProbe(int nx, int ny, int nz) : Grid3D(nx, ny, nz) {}
  void indFrac(double x, double y, double z, int* ind, double* frac) {
    getGridIndAndFrac(x, ind[0], frac[0], y, ind[1], frac[1], z, ind[2], frac[2]);
  }
};
}  // namespace

// Grid2D::binValue for each particle (Grid2D.cc:142-166); rho is [nx][ny].
REF_API void ref_grid2d_bin(int nx, int ny, const double* lims, long n, const double* x,
                            const double* y, const double* m, double* rho) {
  Grid2D g(nx, ny, lims[0], lims[1], lims[2], lims[3]);
  for (long i = 0; i < n; i++) g.binValue(m[i], x[i], y[i]);
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) rho[ix * ny + iy] = g.getValueOnGrid(ix, iy);
}

// Grid2D::getIndAndFracX/Y (Grid2D.cc:337-350).
REF_API void ref_grid2d_indfrac(int nx, int ny, const double* lims, long n, const double* x,
                                const double* y, int* ix, double* fx, int* iy, double* fy) {
  Grid2D g(nx, ny, lims[0], lims[1], lims[2], lims[3]);
  for (long i = 0; i < n; i++) {
    g.getIndAndFracX(x[i], ix[i], fx[i]);
    g.getIndAndFracY(y[i], iy[i], fy[i]);
  }
}

// Grid2D::calcGradient(x,y) and getValue(x,y) over a given grid (Grid2D.cc:81-106, 212-256).
REF_API void ref_grid2d_grad(int nx, int ny, const double* lims, const double* arr, long n,
                             const double* x, const double* y, double* ex, double* ey,
                             double* val) {
  Grid2D g(nx, ny, lims[0], lims[1], lims[2], lims[3]);
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) g.setValue(arr[ix * ny + iy], ix, iy);
  for (long i = 0; i < n; i++) {
    g.calcGradient(x[i], y[i], ex[i], ey[i]);
    val[i] = g.getValue(x[i], y[i]);
  }
}

// Grid1D::binValue for each particle (Grid1D.cc:422-440).
REF_API void ref_grid1d_bin(int nz, double zmin, double zmax, long n, const double* z,
                            const double* m, double* out) {
  Grid1D g(nz, zmin, zmax);
  for (long i = 0; i < n; i++) g.binValue(m[i], z[i]);
  for (int k = 0; k < nz; k++) out[k] = g.getValueOnGrid(k);
}

// Grid1D::getValue and calcGradient over a given line (Grid1D.cc:191-210, 476-489).
REF_API void ref_grid1d_value(int nz, double zmin, double zmax, const double* arr, long n,
                              const double* z, double* val, double* grad) {
  Grid1D g(nz, zmin, zmax);
  for (int k = 0; k < nz; k++) g.setValue(arr[k], k);
  for (long i = 0; i < n; i++) {
    val[i] = g.getValue(z[i]);
    g.calcGradient(z[i], grad[i]);
  }
}

// Grid3D::binValue (Grid3D.cc:307-432); rho is [nz][nx][ny].
REF_API void ref_grid3d_bin(int nx, int ny, int nz, const double* lims, long n, const double* x,
                            const double* y, const double* z, const double* m, double* rho) {
  Grid3D g(nx, ny, nz);
  g.setGridX(lims[0], lims[1]);
  g.setGridY(lims[2], lims[3]);
  g.setGridZ(lims[4], lims[5]);
  g.setZero();
  for (long i = 0; i < n; i++) g.binValue(m[i], x[i], y[i], z[i]);
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) rho[((long)iz * nx + ix) * ny + iy] = g.getValueOnGrid(ix, iy, iz);
}

// Grid3D::getGridIndAndFrac (Grid3D.cc:226-264).
REF_API void ref_grid3d_indfrac(int nx, int ny, int nz, const double* lims, long n,
                                const double* x, const double* y, const double* z, int* ind,
                                double* frac) {
  Grid3DProbe g(nx, ny, nz);
  g.setGridX(lims[0], lims[1]);
  g.setGridY(lims[2], lims[3]);
  g.setGridZ(lims[4], lims[5]);
  for (long i = 0; i < n; i++) g.indFrac(x[i], y[i], z[i], ind + 3 * i, frac + 3 * i);
}

// Grid3D::calcGradient over a given grid (Grid3D.cc:438-532, 718-733).
REF_API void ref_grid3d_grad(int nx, int ny, int nz, const double* lims, const double* arr, long n,
                             const double* x, const double* y, const double* z, double* gx,
                             double* gy, double* gz) {
  Grid3D g(nx, ny, nz);
  g.setGridX(lims[0], lims[1]);
  g.setGridY(lims[2], lims[3]);
  g.setGridZ(lims[4], lims[5]);
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) g.setValue(arr[((long)iz * nx + ix) * ny + iy], ix, iy, iz);
  for (long i = 0; i < n; i++) g.calcGradient(x[i], gx[i], y[i], gy[i], z[i], gz[i]);
}

// Grid3D::binBunch(bunch, lambda) with longWrapping = 1 (Grid3D.cc:281-305): z is wrapped
// with remainder(z, lambda) and the z stencil is periodic.
REF_API void ref_grid3d_bin_wrapped(int nx, int ny, int nz, const double* lims, long n,
                                    const double* x, const double* y, const double* z,
                                    const double* m, double lambda, double* rho) {
  Grid3D g(nx, ny, nz);
  g.setGridX(lims[0], lims[1]);
  g.setGridY(lims[2], lims[3]);
  g.setGridZ(lims[4], lims[5]);
  g.setZero();
  g.setLongWrapping(1);
  for (long i = 0; i < n; i++) g.binValue(m[i], x[i], y[i], remainder(z[i], lambda));
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) rho[((long)iz * nx + ix) * ny + iy] = g.getValueOnGrid(ix, iy, iz);
}

// Grid3D::calcGradient with longWrapping = 1.
REF_API void ref_grid3d_grad_wrapped(int nx, int ny, int nz, const double* lims, const double* arr,
                                     long n, const double* x, const double* y, const double* z,
                                     double* gx, double* gy, double* gz) {
  Grid3D g(nx, ny, nz);
  g.setGridX(lims[0], lims[1]);
  g.setGridY(lims[2], lims[3]);
  g.setGridZ(lims[4], lims[5]);
  g.setLongWrapping(1);
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) g.setValue(arr[((long)iz * nx + ix) * ny + iy], ix, iy, iz);
  for (long i = 0; i < n; i++) g.calcGradient(x[i], gx[i], y[i], gy[i], z[i], gz[i]);
}

// Base/ShapedBoundary2D (BaseBoundary2D.cc, ShapedBoundary2D.cc): shape 0 is a
// BaseBoundary2D whose nPoints points come from px, py; 1/2/3 is a ShapedBoundary2D
// Circle/Ellipse/Rectangle of dimensions xDim, yDim.  Writes the boundary points
// (bx, by: nPoints after the shape's rounding to a multiple of 4), the limits
// lims[4], isInside for n test points, and phi after addBoundaryPotential(rho, phi) on a
// [nx][ny] grid spanning glims.  Returns the number of points.
REF_API int ref_boundary2d(int shape, int nPoints, int nModes, double xDim, double yDim,
                           const double* px, const double* py, double* bx, double* by,
                           double* lims, long n, const double* tx, const double* ty,
                           int* inside, int nx, int ny, const double* glims, double* phi) {
  static const char* names[4] = {"", "Circle", "Ellipse", "Rectangle"};
  BaseBoundary2D* b;
  if (shape == 0) {
    b = new BaseBoundary2D(nPoints, nModes);
    for (int i = 0; i < nPoints; i++) b->setBoundaryPoint(i, px[i], py[i]);
    b->initializeBPs();
  } else {
    b = new ShapedBoundary2D(nPoints, nModes, std::string(names[shape]), xDim, yDim);
  }
  const int np = b->getNumberOfPoints();
  for (int i = 0; i < np; i++) {
    bx[i] = b->getBoundaryPointX(i);
    by[i] = b->getBoundaryPointY(i);
  }
  lims[0] = b->getMinX(); lims[1] = b->getMaxX(); lims[2] = b->getMinY(); lims[3] = b->getMaxY();
  for (long i = 0; i < n; i++) inside[i] = b->isInside(tx[i], ty[i]);
  if (phi) {
    Grid2D rho(nx, ny, glims[0], glims[1], glims[2], glims[3]);
    Grid2D ph(nx, ny, glims[0], glims[1], glims[2], glims[3]);
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) ph.setValue(phi[ix * ny + iy], ix, iy);
    b->addBoundaryPotential(&rho, &ph);
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) phi[ix * ny + iy] = ph.getValueOnGrid(ix, iy);
  }
  delete b;
  return np;
}
