/*
 * orbit_oracle.h -- CPU restatement of the py-orbit tracking + space-charge path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * MI355X implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never calls it.
 *
 * Every function restates one reference routine with the same formulas,
 * evaluation order and quirks (cited file:line, paths relative to the
 * reference tree root).  Coordinates are FP64 SoA arrays owned by the caller.
 */
#ifndef ORBIT_ORACLE_H
#define ORBIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bunch attributes + synchronous particle (src/orbit/Bunch.cc:404-425,
 * src/orbit/SyncPart.cc:236-317). */
typedef struct {
  double mass;             /* GeV */
  double charge;           /* e */
  double classical_radius; /* m */
  double macrosize;        /* bunch macroSize (used when msize == NULL) */
  double T, p, beta, gamma;/* synchronous particle kinematics */
  double time;             /* synchronous particle time, s */
} orc_sync;

typedef struct {
  int64_t n;
  double *x, *xp, *y, *yp, *z, *dE;
  const double *msize; /* optional per-particle "macrosize" attribute */
  orc_sync s;
} orc_bunch;

void orc_set_threads(int nthreads);
int orc_get_threads(void);

/* SyncPart kinEnergy(T) setter: wrap_syncpart.cc:70-92, SyncPart.cc:133-138,
 * 236-317. */
void orc_sync_set_kin_energy(orc_sync* s, double T);

/* ---- teapot_base maps (src/teapot/teapotbase.cc) ---- */
void orc_rotatexy(orc_bunch* b, double anglexy);
void orc_drift(orc_bunch* b, double length);
void orc_wrapbunch(orc_bunch* b, double length);
void orc_kick(orc_bunch* b, double kx, double ky, double kE);
void orc_multp(orc_bunch* b, int pole, double kl, int skew);
void orc_multpfringeIN(orc_bunch* b, int pole, double kl, int skew);
void orc_multpfringeOUT(orc_bunch* b, int pole, double kl, int skew);
void orc_quad1(orc_bunch* b, double length, double kq);
void orc_quad2(orc_bunch* b, double length);
void orc_quadfringeIN(orc_bunch* b, double kq);
void orc_quadfringeOUT(orc_bunch* b, double kq);
/* bends, wedges, solenoid, ring RF (teapotbase.cc:923-1569) */
void orc_wedgerotate(orc_bunch* b, double e, int frinout);
void orc_wedgedrift(orc_bunch* b, double e, int inout);
void orc_wedgebend(orc_bunch* b, double e, int inout, double rho, int nsteps);
void orc_wedgebendCF(orc_bunch* b, double e, int inout, double rho, int vecnum, const int* pole,
                     const double* kl, const int* skew, int nsteps);
void orc_bend1(orc_bunch* b, double length, double th);
void orc_bend2(orc_bunch* b, double length);
void orc_bend3(orc_bunch* b, double th);
void orc_bend4(orc_bunch* b, double th);
void orc_bendfringeIN(orc_bunch* b, double rho);
void orc_bendfringeOUT(orc_bunch* b, double rho);
void orc_soln(orc_bunch* b, double length, double B);
void orc_ringrf(orc_bunch* b, double ring_length, int harmonic_numb, double voltage,
                double phase_s, int useCharge);

/* ---- extrema (src/utils/bunch/BunchExtremaCalculator.cc:37-83) ---- */
void orc_extrema_xyz(const orc_bunch* b, double out6[6]);
void orc_extrema6(const orc_bunch* b, double out12[12]);
double orc_extrema_r(const orc_bunch* b);
/* ---- computeBunchMoments (src/orbit/BunchDiagnostics/BunchTwissAnalysis.cc:145-337) ---- */
void orc_compute_bunch_moments(const orc_bunch* b, int order, int dispflag, double xAvg,
                               double yAvg, double D, double Em, double b2, double xb,
                               double yb, double* mom);

/* ---- Grid2D (src/spacecharge/Grid2D.cc) ----  arr is row-major [nx][ny] */
typedef struct {
  int nx, ny;
  double xMin, xMax, yMin, yMax, dx, dy;
  double* arr;
} orc_grid2d;
void orc_grid2d_set_x(orc_grid2d* g, double xMin, double xMax);
void orc_grid2d_set_y(orc_grid2d* g, double yMin, double yMax);
void orc_grid2d_ind_frac_x(const orc_grid2d* g, double x, int* ind, double* frac);
void orc_grid2d_ind_frac_y(const orc_grid2d* g, double y, int* ind, double* frac);
void orc_grid2d_bin_value(orc_grid2d* g, double value, double x, double y);
void orc_grid2d_bin_bunch(orc_grid2d* g, const orc_bunch* b);
void orc_grid2d_gradient(const orc_grid2d* g, double x, double y, double* ex, double* ey);
double orc_grid2d_value(const orc_grid2d* g, double x, double y);

/* ---- Grid1D (src/spacecharge/Grid1D.cc) ---- cell-centred, periodic */
typedef struct {
  int nz;
  double zMin, zMax, dz;
  double* arr;
} orc_grid1d;
void orc_grid1d_set_z(orc_grid1d* g, double zMin, double zMax);
void orc_grid1d_ind_wz(const orc_grid1d* g, double z, int* i0, int* ip, double* w0, double* wp);
void orc_grid1d_bin_value(orc_grid1d* g, double value, double z);
void orc_grid1d_bin_bunch(orc_grid1d* g, const orc_bunch* b);
void orc_grid1d_bin_axis(orc_grid1d* g, const orc_bunch* b, int axis, int by_particle);
double orc_grid1d_value(const orc_grid1d* g, double z);
void orc_grid1d_gradient(const orc_grid1d* g, double z, double* ez);

/* ---- Grid3D (src/spacecharge/Grid3D.cc) ---- arr is [nz][nx][ny] */
typedef struct {
  int nx, ny, nz;
  double xMin, xMax, yMin, yMax, zMin, zMax, dx, dy, dz;
  double* arr;
  int wrap; /* longWrapping (Grid3D.cc:187-197): periodic z clamp and stencil */
} orc_grid3d;
void orc_grid3d_set_x(orc_grid3d* g, double a, double b);
void orc_grid3d_set_y(orc_grid3d* g, double a, double b);
void orc_grid3d_set_z(orc_grid3d* g, double a, double b);
void orc_grid3d_ind_frac(const orc_grid3d* g, double x, double y, double z, int ind[3], double frac[3]);
void orc_grid3d_bin_value(orc_grid3d* g, double m, double x, double y, double z);
void orc_grid3d_bin_bunch(orc_grid3d* g, const orc_bunch* b);
/* binBunch(bunch, lambda) with longWrapping = 1 (Grid3D.cc:281-305): z -> remainder(z, lambda) */
void orc_grid3d_bin_bunch_wrapped(orc_grid3d* g, const orc_bunch* b, double lambda);
void orc_grid3d_gradient(const orc_grid3d* g, double x, double y, double z,
                         double* gx, double* gy, double* gz);

/* ---- FFT (stand-in for FFTW's conventions: forward e^{-i}, unnormalised) ---- */
void orc_fft_c2c(double* re, double* im, int n, int inverse);
void orc_fft_r2c_2d(const double* in, int n0, int n1, double* out_cplx /* [n0][n1/2+1][2] */);
void orc_fft_c2r_2d(const double* in_cplx, int n0, int n1, double* out);
void orc_fft_r2c_3d(const double* in, int n0, int n1, int n2, double* out_cplx);
void orc_fft_c2r_3d(const double* in_cplx, int n0, int n1, int n2, double* out);
/* Replace both 3D transforms by caller functions with the same signature and conventions
 * (NULL, NULL restores the oracle's own FFT). */
typedef void (*orc_fft3_fn)(const double* in, int n0, int n1, int n2, double* out);
void orc_set_fft3_hook(orc_fft3_fn r2c, orc_fft3_fn c2r);

/* ---- PoissonSolverFFT2D (src/spacecharge/PoissonSolverFFT2D.cc) ---- */
typedef struct orc_poisson2d orc_poisson2d;
orc_poisson2d* orc_poisson2d_create(int nx, int ny, double xMin, double xMax, double yMin, double yMax);
void orc_poisson2d_destroy(orc_poisson2d* p);
void orc_poisson2d_find_potential(orc_poisson2d* p, const orc_grid2d* rho, orc_grid2d* phi);
const double* orc_poisson2d_green_hat(const orc_poisson2d* p); /* [2nx][ny+1][2] */

/* ---- PoissonSolverFFT3D (src/spacecharge/PoissonSolverFFT3D.cc) ---- */
typedef struct orc_poisson3d orc_poisson3d;
orc_poisson3d* orc_poisson3d_create(int nx, int ny, int nz, double xMin, double xMax,
                                    double yMin, double yMax, double zMin, double zMax);
void orc_poisson3d_destroy(orc_poisson3d* p);
void orc_poisson3d_set_grid(orc_poisson3d* p, double xMin, double xMax, double yMin, double yMax,
                            double zMin, double zMax);
void orc_poisson3d_find_potential(orc_poisson3d* p, const orc_grid3d* rho, orc_grid3d* phi);
void orc_poisson3d_limits(const orc_poisson3d* p, double out6[6]);
/* setNumberOfExternalBunches (odd counts round up, PoissonSolverFFT3D.cc:101-106) and
 * setSpacingOfExternalBunches (:108-110); both take effect at the next Green rebuild. */
void orc_poisson3d_set_ext_bunches(orc_poisson3d* p, int n);
void orc_poisson3d_set_ext_spacing(orc_poisson3d* p, double lambda);

/* ---- SpaceChargeCalc2p5Drb (src/spacecharge/SpaceChargeCalc2p5Drb.cc) ---- */
typedef struct orc_sc2p5drb orc_sc2p5drb;
orc_sc2p5drb* orc_sc2p5drb_create(int nx, int ny, int nz, double xy_ratio);
void orc_sc2p5drb_destroy(orc_sc2p5drb* c);
void orc_sc2p5drb_set_long_avg_n(orc_sc2p5drb* c, int n);
/* returns 0 on success, -1 if the bunch has a zero-size extent (reference aborts) */
int orc_sc2p5drb_track(orc_sc2p5drb* c, orc_bunch* b, double length, double pipe_radius);
/* which: 0 rho, 1 phi, 2 z line density, 3 z derivative.  lims gets the grid limits
 * (x/y: xMin,xMax,yMin,yMax ; z: zMin,zMax).  Returns element count. */
int orc_sc2p5drb_get_grid(const orc_sc2p5drb* c, int which, double* out, double* lims);
/* x_c, y_c, a_bunch, totalMacrosize of the last bunchAnalysis */
void orc_sc2p5drb_get_moments(const orc_sc2p5drb* c, double out4[4]);

/* Base/ShapedBoundary2D (BaseBoundary2D.cc, ShapedBoundary2D.cc); shape 1 circle,
 * 2 ellipse, 3 rectangle. */
typedef struct orc_boundary2d orc_boundary2d;
orc_boundary2d* orc_boundary2d_create(int nPoints, int nModes);
orc_boundary2d* orc_boundary2d_create_shaped(int nPoints, int nModes, int shape, double xDim,
                                             double yDim);
void orc_boundary2d_destroy(orc_boundary2d* b);
void orc_boundary2d_set_point(orc_boundary2d* b, int i, double x, double y);
int orc_boundary2d_initialize(orc_boundary2d* b);
int orc_boundary2d_is_inside(const orc_boundary2d* b, double x, double y);
void orc_boundary2d_add_potential(orc_boundary2d* b, orc_grid2d* phi);
void orc_boundary2d_info(const orc_boundary2d* b, int* np, double lims[4]);
void orc_boundary2d_points(const orc_boundary2d* b, double* x, double* y);

/* SpaceChargeCalc2p5D without a boundary (SpaceChargeCalc2p5D.cc) */
typedef struct orc_sc2p5d orc_sc2p5d;
void orc_poisson2d_set_grid(orc_poisson2d* p, double xMin, double xMax, double yMin, double yMax);
orc_sc2p5d* orc_sc2p5d_create(int nx, int ny, int nz, double xy_ratio);
void orc_sc2p5d_destroy(orc_sc2p5d* c);
int orc_sc2p5d_track(orc_sc2p5d* c, orc_bunch* b, double length);
int orc_sc2p5d_track_boundary(orc_sc2p5d* c, orc_bunch* b, double length, orc_boundary2d* bd);
int orc_sc2p5d_get_grid(const orc_sc2p5d* c, int which, double* out, double* lims);
void orc_sc2p5d_state(const orc_sc2p5d* c, double out7[7]);

/* ---- SpaceChargeCalc3D (src/spacecharge/SpaceChargeCalc3D.cc) ---- */
/* SpaceChargeCalcSliceBySlice2D (SpaceChargeCalcSliceBySlice2D.cc); bd may be NULL */
typedef struct orc_slice2d orc_slice2d;
orc_slice2d* orc_slice2d_create(int nx, int ny, int nz, double xy_ratio);
void orc_slice2d_destroy(orc_slice2d* c);
int orc_slice2d_track(orc_slice2d* c, orc_bunch* b, double length, orc_boundary2d* bd);
void orc_slice2d_get_grid(const orc_slice2d* c, int which, double* out, double* lims);
void orc_slice2d_state(const orc_slice2d* c, double out6[6]);

typedef struct orc_sc3d orc_sc3d;
orc_sc3d* orc_sc3d_create(int nx, int ny, int nz);
void orc_sc3d_destroy(orc_sc3d* c);
void orc_sc3d_set_ratio_limit(orc_sc3d* c, double r);
int orc_sc3d_track(orc_sc3d* c, orc_bunch* b, double length);
/* which: 0 rho, 1 phi; lims: xMin,xMax,yMin,yMax,zMin,zMax (rest frame z) */
int64_t orc_sc3d_get_grid(const orc_sc3d* c, int which, double* out, double* lims);
int orc_sc3d_green_rebuilds(const orc_sc3d* c);
/* numExtBunches / freqOfBunches (SpaceChargeCalc3D.cc:84-97); n > 0 selects
 * wrappedBunchAnalysis (:251-328) and the periodic z kick (:138-140). */
void orc_sc3d_set_num_ext_bunches(orc_sc3d* c, int n);
int orc_sc3d_get_num_ext_bunches(const orc_sc3d* c);
void orc_sc3d_set_freq(orc_sc3d* c, double f);
double orc_sc3d_get_freq(const orc_sc3d* c);

/* ---- BunchTwissAnalysis emittance (src/orbit/BunchDiagnostics/BunchTwissAnalysis.cc) ---- */
void orc_twiss_emittances(const orc_bunch* b, double out3[3]);


/* ---- flat-array entry points (same signatures as oracle/ref_harness.cc) ---- */
void orc_flat_grid2d_bin(int nx, int ny, const double* l, long n, const double* x, const double* y,
                         const double* m, double* rho);
void orc_flat_grid2d_indfrac(int nx, int ny, const double* l, long n, const double* x,
                             const double* y, int* ix, double* fx, int* iy, double* fy);
void orc_flat_grid2d_grad(int nx, int ny, const double* l, const double* arr, long n,
                          const double* x, const double* y, double* ex, double* ey, double* val);
void orc_flat_bunch_bin_axis(const orc_bunch* b, int nz, double zmin, double zmax, int axis,
                             int by_particle, double* out);
void orc_flat_grid1d_bin(int nz, double zmin, double zmax, long n, const double* z, const double* m,
                         double* out);
void orc_flat_grid1d_value(int nz, double zmin, double zmax, const double* arr, long n,
                           const double* z, double* val, double* grad);
/* longWrapping = 1 variants: bin with z -> remainder(z, lambda); gradient with periodic z */
void orc_flat_grid3d_bin_wrapped(int nx, int ny, int nz, const double* l, long n, const double* x,
                                 const double* y, const double* z, const double* m, double lambda,
                                 double* rho);
void orc_flat_grid3d_grad_wrapped(int nx, int ny, int nz, const double* l, const double* arr,
                                  long n, const double* x, const double* y, const double* z,
                                  double* gx, double* gy, double* gz);
/* 3D solve with external-bunch images in the Green function */
void orc_flat_poisson3d_images(int nx, int ny, int nz, const double* sl, const double* rl,
                               int nbunches, double lambda, const double* rho, double* phi);
void orc_flat_grid3d_bin(int nx, int ny, int nz, const double* l, long n, const double* x,
                         const double* y, const double* z, const double* m, double* rho);
void orc_flat_grid3d_indfrac(int nx, int ny, int nz, const double* l, long n, const double* x,
                             const double* y, const double* z, int* ind, double* frac);
void orc_flat_grid3d_grad(int nx, int ny, int nz, const double* l, const double* arr, long n,
                          const double* x, const double* y, const double* z, double* gx,
                          double* gy, double* gz);
void orc_flat_poisson2d(int nx, int ny, const double* solver_lims, const double* rho_lims,
                        const double* rho, double* phi);
void orc_flat_poisson3d(int nx, int ny, int nz, const double* solver_lims,
                        const double* rho_lims, const double* rho, double* phi);
void orc_flat_long_derivative(int nz, double zmin, double zmax, int n_avg, const double* zg,
                              double* zd);

#ifdef __cplusplus
}
#endif
#endif
