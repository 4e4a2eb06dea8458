/*
 * orbit_mi355x.h -- C ABI of liborbit_mi355x.so, the MI355X (gfx950) implementation
 * of py-orbit's TEAPOT macroparticle tracking and space-charge kick.
 *
 * Every entry point returns 0 on success and a non-zero status on failure; the
 * message of the last failure on the calling thread is ob_last_error().  (The
 * reference aborts the process instead: src/mpi/orbit_mpi.cc:37-75.)
 *
 * Data stays on the device.  Calls are asynchronous on the library's stream:
 * host reads (ob_bunch_get_*) and ob_synchronize() wait for the work queued
 * before them and report any error a device kernel flagged meanwhile.
 * Handles are not thread-safe.
 *
 * Reference interfaces replaced (paths relative to the reference tree root):
 *   bunch        src/orbit/wrap_bunch.cc:1178-1232, wrap_syncpart.cc:549-566
 *   teapot_base  src/teapot/wrap_teapotbase.cc:488-512 (arg formats :29-473)
 *   spacecharge  src/spacecharge/wrap_spacechargecalc2p5d_rb.cc:38-51,189-198,
 *                wrap_spacechargecalc3d.cc:38-45,171-177
 */
#ifndef ORBIT_MI355X_H
#define ORBIT_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ob_bunch ob_bunch;
typedef struct ob_sc2p5drb ob_sc2p5drb;
typedef struct ob_sc2p5d ob_sc2p5d;
typedef struct ob_boundary2d ob_boundary2d;
typedef struct ob_slice2d ob_slice2d;
typedef struct ob_sc3d ob_sc3d;

/* ---- library ---------------------------------------------------------- */
/* Select the device and create the library stream.  Called implicitly by the
 * first create call with device 0.  (src/main/main.cc:27-82 starts the runtime.) */
int ob_init(int device);
/* Message of the last failure on this thread ("" if none). */
const char* ob_last_error(void);
/* Wait for all queued work; returns non-zero if a kernel flagged an error
 * (e.g. a zero-extent bunch in a space-charge calculator). */
int ob_synchronize(void);
/* Library version string. */
const char* ob_version(void);

/* ---- Bunch (src/orbit/Bunch.hh:54-293, wrap_bunch.cc:1178-1232) --------- */
/* Bunch(): protons, macroSize 0, synchronous kinetic energy 0 (Bunch.cc:404-425). */
int ob_bunch_create(ob_bunch** out);
int ob_bunch_destroy(ob_bunch* b);
/* mass(v) GeV, charge(v) e, classicalRadius, macroSize(v): wrap_bunch.cc "mass",
 * "charge", "classicalRadius", "macroSize". */
int ob_bunch_set_mass(ob_bunch* b, double mass);
int ob_bunch_set_charge(ob_bunch* b, double charge);
int ob_bunch_set_classical_radius(ob_bunch* b, double r0);
int ob_bunch_set_macrosize(ob_bunch* b, double macrosize);
/* out[4] = mass, charge, classicalRadius, macroSize */
int ob_bunch_get_attrs(ob_bunch* b, double* out);
/* SyncPart.kinEnergy(T) setter (wrap_syncpart.cc:70-92). */
int ob_bunch_set_kin_energy(ob_bunch* b, double T_GeV);
/* SyncPart.momentum(p) setter (wrap_syncpart.cc:136-156, SyncPart.cc:133-138, 236-317). */
int ob_bunch_set_sync_momentum(ob_bunch* b, double p_GeV);
/* out[6] = kinEnergy, momentum, beta, gamma, time, mass (wrap_syncpart.cc:549-566) */
int ob_bunch_get_sync(ob_bunch* b, double* out);
int ob_bunch_set_sync_time(ob_bunch* b, double t);
/* addParticle(x,xp,y,yp,z,dE) (wrap_bunch.cc "addParticle"). */
int ob_bunch_add_particle(ob_bunch* b, double x, double xp, double y, double yp, double z,
                          double dE);
/* Bulk replacement of all particles from host SoA arrays (n each). */
int ob_bunch_set_coords(ob_bunch* b, int64_t n, const double* x, const double* xp,
                        const double* y, const double* yp, const double* z, const double* dE);
/* Copy all coordinates to host SoA arrays (each of getSize() doubles). */
int ob_bunch_get_coords(ob_bunch* b, double* x, double* xp, double* y, double* yp, double* z,
                        double* dE);
/* One coordinate: k = 0..5 for x, xp, y, yp, z, dE (wrap_bunch.cc "x".."dE"). */
int ob_bunch_get_coord(ob_bunch* b, int64_t i, int k, double* v);
int ob_bunch_set_coord(ob_bunch* b, int64_t i, int k, double v);
/* getSize / getSizeGlobal (one process: identical, Bunch.cc:850-861). */
int ob_bunch_get_size(ob_bunch* b, int64_t* n);
/* deleteParticle(i) marks the particle dead; compress() removes dead particles
 * keeping order (Bunch.cc:704-763). */
int ob_bunch_delete_particle(ob_bunch* b, int64_t i);
int ob_bunch_compress(ob_bunch* b);
/* Per-particle "macrosize" attribute (ParticleMacroSize; read at Grid2D.cc:122-131).
 * n must equal getSize(); NULL removes the attribute. */
int ob_bunch_set_macrosizes(ob_bunch* b, int64_t n, const double* m);
int ob_bunch_get_macrosizes(ob_bunch* b, int64_t n, double* m);
/* The sums BunchTwissAnalysis::analyzeBunch accumulates (BunchTwissAnalysis.cc:49-142),
 * on the device after the queued maps: out[0] = sum w, out[1+k] = sum w c_k (k = 0..5),
 * out[7..27] = sum w c_a c_b for a <= b row by row (a = 0: b = 0..5, a = 1: b = 1..5 ...),
 * w = the macrosize attribute or 1.  Dead particles are counted until compress(), as the
 * reference counts them until its own compress() call. */
int ob_bunch_moments(ob_bunch* b, double* out);
/* ob_bunch_moments without any device allocation (no counterpart in the reference): for
 * callers that cannot allocate (code under HIP graph capture, tight per-turn monitoring
 * loops, harnesses that forbid hipMalloc).  Same kernels and the same out[28] bit for bit;
 * the partial sums use a fixed per-device scratch of 229,600 bytes that the library holds
 * until the process ends, which fits every bunch size (at most 1024 workgroup partials).
 * Calls that overlap on one device race on that scratch.  Refuses a bunch that belongs to
 * a batch (use ob_bunch_moments). */
int ob_bunch_moments_noalloc(ob_bunch* b, double* out);
/* BunchExtremaCalculator::getExtremaXYZ and getExtremaXpYpdE
 * (src/utils/bunch/BunchExtremaCalculator.cc:37-120, wrap_bunch_extrema_calculator.cc) in
 * one device pass: out[12] = (min, max) of x, xp, y, yp, z, dE; DBL_MAX, -DBL_MAX for an
 * empty bunch. */
int ob_bunch_extrema(ob_bunch* b, double* out);
/* BunchExtremaCalculator::getExtremaR (BunchExtremaCalculator.cc:167-208,
 * wrap_bunch_extrema_calculator.cc:95-109): *rmax = the largest sqrt((x - <x>)^2 +
 * (y - <y>)^2) over the particles, <.> the unweighted mean; 0 for an empty bunch. */
int ob_bunch_extrema_r(ob_bunch* b, double* rmax);
/* Grid1D::binBunch(bunch, axis) and binBunchByParticle(bunch, axis)
 * (src/spacecharge/Grid1D.cc:241-275, 335-352; wrap_grid1D.cc:244-325) on the device, for a
 * grid of nz cells (1 <= nz <= 2^26) over [zMin, zMax]: the binned values are added to the
 * host array arr[nz].  by_particle != 0 weighs every particle 1. */
int ob_grid1d_bin_bunch(ob_bunch* b, int axis, int by_particle, int nz, double zMin,
                        double zMax, double* arr);
/* The particle sums of BunchTwissAnalysis::computeBunchMoments
 * (BunchTwissAnalysis.cc:145-337), 1 <= order <= 12.  prm[8] = <x>, <y>, D_x,
 * E_kin + mass, beta^2, xb, yb, dispersion flag; out[order (order + 3) / 2 + 2] = the sums
 * for j = 0..order-1, i = 0..order-j, then the stray [2][order] entry of a bunch without
 * macrosizes (order >= 2), then the total weight. */
int ob_bunch_xy_moments(ob_bunch* b, int order, const double* prm, double* out);
/* Device pointers of the SoA coordinate arrays (x, xp, y, yp, z, dE), valid until the
 * next size change; pending maps are applied first. */
int ob_bunch_device_coords(ob_bunch* b, double** ptrs6);

/* ---- teapot_base maps (src/teapot/teapotbase.cc; wrap_teapotbase.cc:488-512) ----
 * Maps are queued on the bunch and applied by one fused kernel when the bunch is
 * next needed (a space-charge kick, a host read, ob_bunch_flush); results are
 * identical to applying them one by one.  useCharge is accepted and ignored as in
 * the reference (teapotbase.cc:620-700, 336-372). */
int ob_tpb_rotatexy(ob_bunch* b, double angle);                        /* :80   */
int ob_tpb_drift(ob_bunch* b, double length);                          /* :158  */
int ob_tpb_wrapbunch(ob_bunch* b, double length);                      /* :202  */
int ob_tpb_kick(ob_bunch* b, double kx, double ky, double kE, int useCharge); /* :232 */
int ob_tpb_multp(ob_bunch* b, int pole, double kl, int skew, int useCharge);  /* :336 */
int ob_tpb_multpfringeIN(ob_bunch* b, int pole, double kl, int skew, int useCharge);  /* :395 */
int ob_tpb_multpfringeOUT(ob_bunch* b, int pole, double kl, int skew, int useCharge); /* :508 */
int ob_tpb_quad1(ob_bunch* b, double length, double kq, int useCharge); /* :620 */
int ob_tpb_quad2(ob_bunch* b, double length);                           /* :719 */
int ob_tpb_quadfringeIN(ob_bunch* b, double kq, int useCharge);         /* :784 */
int ob_tpb_quadfringeOUT(ob_bunch* b, double kq, int useCharge);        /* :854 */
int ob_tpb_quad3(ob_bunch* b, double length, double kq, int useCharge); /* :761 (no-op) */
/* drifti(bunch, i, length) (:120): one particle, through the host copy (slow). */
int ob_tpb_drifti(ob_bunch* b, int64_t i, double length);
/* Bends, wedges, solenoid and ring RF (:923-1566).  bend1 and soln advance the
 * synchronous time as drift does; RingRF with phase_s != 0 changes the synchronous
 * momentum.  wedgebendCF takes n <= 8 multipole terms with poles 0..15 (the device op
 * carries them inline). */
int ob_tpb_wedgerotate(ob_bunch* b, double e, int frinout);             /* :923  */
int ob_tpb_wedgedrift(ob_bunch* b, double e, int inout);                /* :986  */
int ob_tpb_wedgebend(ob_bunch* b, double e, int inout, double rho, int nsteps); /* :1037 */
int ob_tpb_bend1(ob_bunch* b, double length, double th);                /* :1099 */
int ob_tpb_bend2(ob_bunch* b, double length);                           /* :1164 */
int ob_tpb_bend3(ob_bunch* b, double th);                               /* :1208 */
int ob_tpb_bend4(ob_bunch* b, double th);                               /* :1251 */
int ob_tpb_bendfringeIN(ob_bunch* b, double rho);                       /* :1292 */
int ob_tpb_bendfringeOUT(ob_bunch* b, double rho);                      /* :1329 */
int ob_tpb_soln(ob_bunch* b, double length, double B, int useCharge);   /* :1367 */
int ob_tpb_wedgebendCF(ob_bunch* b, double e, int inout, double rho, int n, const int* poles,
                       const double* kls, const int* skews, int nsteps, int useCharge); /* :1447 */
int ob_tpb_RingRF(ob_bunch* b, double ring_length, int harmonic_numb, double voltage,
                  double phase_s, int useCharge);                       /* :1527 */
/* Apply the queued maps now. */
int ob_bunch_flush(ob_bunch* b);

/* A map sequence in one call (the per-SC-interval sequence the lattice layer builds from
 * the TEAPOT nodes, teapot.py:548-976): identical results to calling the entry points
 * above one by one, applied by one fused launch before returning (asynchronously on the
 * library stream).  Arguments per kind (a, b, c real; ia, ib, ic integer):
 *   ROTATEXY a=angle | DRIFT a=length | WRAPBUNCH a=ring length | KICK a,b,c=kx,ky,kE ic=useCharge
 *   MULTP / MULTPFRINGE_IN / _OUT  ia=pole a=kl ib=skew ic=useCharge
 *   QUAD1 a=length b=kq ic=useCharge | QUAD2 a=length | QUADFRINGE_IN / _OUT b=kq ic=useCharge
 *   BEND1 a=length b=th | BEND2 a=length | BEND3 / BEND4 a=th | BENDFRINGE_IN / _OUT a=rho
 *   WEDGEROTATE a=e ia=frinout | WEDGEDRIFT a=e ia=inout | WEDGEBEND a=e ia=inout b=rho ib=nsteps
 *   WEDGEBENDCF as WEDGEBEND plus n, poles, kls, skews, ic=useCharge
 *   SOLN a=length b=B ic=useCharge | RINGRF a=ring length ia=harmonic b=voltage c=phase_s
 *   ic=useCharge */
enum {
  OB_MAP_ROTATEXY = 1, OB_MAP_DRIFT, OB_MAP_WRAPBUNCH, OB_MAP_KICK, OB_MAP_MULTP,
  OB_MAP_MULTPFRINGE_IN, OB_MAP_MULTPFRINGE_OUT, OB_MAP_QUAD1, OB_MAP_QUAD2,
  OB_MAP_QUADFRINGE_IN, OB_MAP_QUADFRINGE_OUT, OB_MAP_BEND1, OB_MAP_BEND2, OB_MAP_BEND3,
  OB_MAP_BEND4, OB_MAP_BENDFRINGE_IN, OB_MAP_BENDFRINGE_OUT, OB_MAP_WEDGEROTATE,
  OB_MAP_WEDGEDRIFT, OB_MAP_WEDGEBEND, OB_MAP_WEDGEBENDCF, OB_MAP_SOLN, OB_MAP_RINGRF
};
typedef struct ob_map_op {
  int32_t kind;
  int32_t ia, ib, ic;
  double a, b, c;
  int32_t n, pad;       /* WEDGEBENDCF: number of multipole terms */
  const int32_t* poles; /* WEDGEBENDCF: n poles, kls, skews */
  const double* kls;
  const int32_t* skews;
} ob_map_op;
int ob_track_sequence(ob_bunch* b, const ob_map_op* ops, int nops);

/* ---- SpaceChargeCalc2p5Drb (src/spacecharge/SpaceChargeCalc2p5Drb.cc) ---- */
/* SpaceChargeCalc2p5Drb(nx, ny, nz[, xy_ratio]) (:31-63). */
int ob_sc2p5drb_create(int nx, int ny, int nz, double xy_ratio, ob_sc2p5drb** out);
int ob_sc2p5drb_destroy(ob_sc2p5drb* c);
/* trackBunch(bunch, length, pipe_radius) (:110-168). */
int ob_sc2p5drb_track(ob_sc2p5drb* c, ob_bunch* b, double length, double pipe_radius);
/* PoissonSolverFFT2D(xSize, ySize, -xy_ratio, xy_ratio, -1, 1).findPotential(rho, phi)
 * (PoissonSolverFFT2D.cc:174-256) with the calculator's solver: rho and phi are host
 * [nx][ny] grids spanning [xMin, xMax] x [yMin, yMax] (same step ratio as the solver). */
int ob_sc2p5drb_find_potential(ob_sc2p5drb* calc, const double* rho, double xMin, double xMax,
                               double yMin, double yMax, double* phi);
/* getRhoGrid / getPhiGrid / getLongGrid / getLongDerivativeGrid (wrap :54-72 and
 * following): which = 0 rho, 1 phi ([nx][ny] row-major), 2 z line density, 3 its
 * derivative ([nz]).  lims receives xMin,xMax,yMin,yMax (which 0,1) or zMin,zMax. */
int ob_sc2p5drb_get_grid(ob_sc2p5drb* c, int which, double* out, double* lims);
/* out[4] = x_center, y_center, a_bunch, totalMacrosize of the last trackBunch. */
int ob_sc2p5drb_get_moments(ob_sc2p5drb* c, double* out);
/* set/getLongAveragingPointsN (:313-320). */
int ob_sc2p5drb_set_long_avg_n(ob_sc2p5drb* c, int n);
int ob_sc2p5drb_get_long_avg_n(ob_sc2p5drb* c, int* n);

/* ---- SpaceChargeCalc2p5D (src/spacecharge/SpaceChargeCalc2p5D.cc), no boundary ---- */
/* SpaceChargeCalc2p5D(nx, ny, nz[, xy_ratio]) (:25-43; wrap_spacechargecalc2p5d.cc:41). */
int ob_sc2p5d_create(int nx, int ny, int nz, double xy_ratio, ob_sc2p5d** out);
int ob_sc2p5d_destroy(ob_sc2p5d* c);
/* trackBunch(bunch, length) (:77-120, bunchAnalysis :122-234): the grid keeps the
 * calculator's x/y ratio while the beam's is within 25% of it, else follows the beam and
 * the solver's Green function is rebuilt on the device; transverse kick only, queued on
 * the bunch like every kick. */
int ob_sc2p5d_track(ob_sc2p5d* c, ob_bunch* b, double length);
/* trackBunch(bunch, length, boundary) (:77-120, 174-234; wrap :139): the grid is the
 * boundary's box, the boundary's potential (BaseBoundary2D.cc:349-458) is added to phi on
 * the device, and particles outside the boundary are not kicked.  bnd NULL = no boundary. */
int ob_sc2p5d_track_boundary(ob_sc2p5d* c, ob_bunch* b, double length, ob_boundary2d* bnd);
/* getRhoGrid / getPhiGrid / getLongGrid (wrap :174-177): which = 0 rho, 1 phi, 2 z line
 * density; lims as for ob_sc2p5drb_get_grid. */
int ob_sc2p5d_get_grid(ob_sc2p5d* c, int which, double* out, double* lims);
/* out[7] = xy_ratio, solver xMin, xMax, yMin, yMax, totalMacrosize of the last step,
 * Green function rebuilds so far (the calculator's private state, :152-172, 206-216). */
int ob_sc2p5d_get_state(ob_sc2p5d* c, double* out);

/* ---- SpaceChargeCalcSliceBySlice2D (src/spacecharge/SpaceChargeCalcSliceBySlice2D.cc;
 *      wrap_spacechargecalc_slicebyslice_2D.cc:41, 107-155) ---- */
/* SpaceChargeCalcSliceBySlice2D(nx, ny, nz[, xy_ratio]) (:15-32). */
int ob_slice2d_create(int nx, int ny, int nz, double xy_ratio, ob_slice2d** out);
int ob_slice2d_destroy(ob_slice2d* c);
/* trackBunch(bunch, length[, boundary]) (:66-113, bunchAnalysis :116-218): 3D TSC deposit,
 * one 2D Poisson solve per z slice (and the boundary's potential per slice), transverse
 * kick xp -= ex*factor, yp -= ey*factor from the 3D gradient of phi; bnd NULL = none. */
int ob_slice2d_track(ob_slice2d* c, ob_bunch* b, double length, ob_boundary2d* bnd);
/* getRhoGrid / getPhiGrid: which 0 rho, 1 phi as [nz][nx][ny]; lims[6] x, y, z limits. */
int ob_slice2d_get_grid(ob_slice2d* c, int which, double* out, double* lims);
/* out[6] = xy_ratio, solver xMin, xMax, yMin, yMax, Green function rebuilds. */
int ob_slice2d_get_state(ob_slice2d* c, double* out);

/* ---- Boundary2D (src/spacecharge/BaseBoundary2D.cc, ShapedBoundary2D.cc;
 *      wrap_boundary2d.cc:38-75, 208-220) ---- */
/* Boundary2D(nPoints, nModes): points set one by one, then initialize (2*nModes+1 must
 * be below nPoints, BaseBoundary2D.cc:57-70). */
int ob_boundary2d_create(int nPoints, int nModes, ob_boundary2d** out);
/* Boundary2D(nPoints, nModes, shape, xDim, yDim): shape "Circle", "Ellipse" or
 * "Rectangle" (ShapedBoundary2D.cc:9-127); initialized on return. */
int ob_boundary2d_create_shaped(int nPoints, int nModes, const char* shape, double xDim,
                                double yDim, ob_boundary2d** out);
int ob_boundary2d_destroy(ob_boundary2d* b);
int ob_boundary2d_set_point(ob_boundary2d* b, int i, double x, double y);   /* :160 */
int ob_boundary2d_get_point(ob_boundary2d* b, int i, double* x, double* y); /* :150 */
/* initialize (initializeBPs, :188-230): limits and the least-squares projector. */
int ob_boundary2d_initialize(ob_boundary2d* b);
/* lims[4] = xMin, xMax, yMin, yMax; info[4] = nPoints, nModes, initialized, shape type
 * (0 no shape, 1 Circle, 2 Ellipse, 3 Rectangle). */
int ob_boundary2d_info(ob_boundary2d* b, double* lims, int* info);
/* isInside(x, y): *out = +1 inside, -1 outside. */
int ob_boundary2d_is_inside(ob_boundary2d* b, double x, double y, int* out);
/* addBoundaryPotential(rhoGrid, phiGrid) (:349-458) on a host phi grid [nx][ny] spanning
 * lims[4] (computed on the device). */
int ob_boundary2d_add_potential(ob_boundary2d* b, double* phi, int nx, int ny,
                                const double* lims);

/* ---- SpaceChargeCalc3D (src/spacecharge/SpaceChargeCalc3D.cc) ---- */
int ob_sc3d_create(int nx, int ny, int nz, ob_sc3d** out);             /* :32-58 */
int ob_sc3d_destroy(ob_sc3d* c);
/* trackBunch(bunch, length) (:100-148).  Blocks until the bunch's queued maps and its
 * extrema are done: the grid sizing and the Green-function rebuild decision
 * (bunchAnalysis, :150-249) run on the host; the rest is queued asynchronously. */
int ob_sc3d_track(ob_sc3d* c, ob_bunch* b, double length);
/* which = 0 rho, 1 phi ([nz][nx][ny]); lims = xMin,xMax,yMin,yMax,zMin,zMax. */
int ob_sc3d_get_grid(ob_sc3d* c, int which, double* out, double* lims);
int ob_sc3d_set_ratio_limit(ob_sc3d* c, double r);                     /* :334-337 */
int ob_sc3d_get_ratio_limit(ob_sc3d* c, double* r);
/* numExtBunches(n) / freqOfBunches(f) (:84-97; wrap_spacechargecalc3d.cc).  n > 0 selects
 * wrappedBunchAnalysis (:251-328): the bunch is binned with z wrapped to one bunch spacing
 * c*beta/f, the Green function holds n image bunches (odd n rounds up, PoissonSolverFFT3D.cc:
 * 101-106) and is rebuilt on every call, and the kick wraps z (:138-140).  Default f = 402.5 MHz. */
int ob_sc3d_set_num_ext_bunches(ob_sc3d* c, int n);
int ob_sc3d_get_num_ext_bunches(ob_sc3d* c, int* n);
int ob_sc3d_set_freq(ob_sc3d* c, double f);
int ob_sc3d_get_freq(ob_sc3d* c, double* f);
/* Number of Green-function rebuilds so far (ratio changes, :186-203). */
int ob_sc3d_green_rebuilds(ob_sc3d* c, int* n);
/* Deposit strategy (no reference counterpart; results agree to FP64 summation order):
 * tiled != 0 sorts the device bunch by xy tile and accumulates each tile in LDS (default),
 * tiled == 0 scatters in particle order with global atomics.  lds_budget_bytes bounds the
 * per-tile LDS block, 1 KiB..116 KiB (0 keeps the current value; default 16 KiB). */
int ob_sc3d_set_deposit_mode(ob_sc3d* c, int tiled, int lds_budget_bytes);

/* ---- batches of bunches (no reference counterpart) ---------------------------------
 * B independent bunches of uneven sizes, each with its own SpaceChargeCalc2p5Drb (grids
 * may differ), sharing one lattice and one synchronous particle (same mass, charge,
 * classical radius and momentum: checked).  Every phase of a step runs as one launch over
 * all bunches; results equal tracking each bunch alone.  The batch owns the bunches' map
 * queues while it is in use: queue maps with ob_batch_track_sequence, not ob_tpb_*. */
typedef struct ob_batch ob_batch;
int ob_batch_create(ob_bunch** bunches, int nb, ob_batch** out);
int ob_batch_destroy(ob_batch* batch);
/* One calculator per bunch, in bunch order. */
int ob_batch_set_sc2p5drb(ob_batch* batch, ob_sc2p5drb** calcs);
/* Queue maps for every bunch (applied lazily, like ob_tpb_*). */
int ob_batch_track_sequence(ob_batch* batch, const ob_map_op* ops, int nops);
/* SpaceChargeCalc2p5Drb.trackBunch(bunch, length, pipe_radius) of every bunch. */
int ob_batch_sc2p5drb_track(ob_batch* batch, double length, double pipe_radius);
int ob_batch_flush(ob_batch* batch);
/* Wait; fails if a bunch had zero extent in its last SC step (the reference aborts). */
int ob_batch_synchronize(ob_batch* batch);
/* Autotuning: particles per push/binning work item (chunk) and per tile-deposit work item
 * (deposit_item).  Tuned by timing on the first SC step unless set here; the choice is kept
 * for the process per device and batch shape (bunch sizes, grids), and a batch of a shape
 * already tuned takes it without timing. */
int ob_batch_set_tuning(ob_batch* batch, int chunk, int deposit_item);
int ob_batch_get_tuning(ob_batch* batch, int* chunk, int* deposit_item, int* tuned);
/* hipGraph: the batch calls between begin and end are recorded (not run) and replayed by
 * ob_batch_replay(count).  Pending work is applied before and at the end of the capture,
 * so a replay starts and ends with nothing pending.  Bunch sizes must not change between
 * the capture and its replays. */
int ob_batch_capture_begin(ob_batch* batch);
int ob_batch_capture_end(ob_batch* batch);
int ob_batch_replay(ob_batch* batch, int count);

/* ---- learned denoiser of deposited charge density (no reference counterpart) ----
 * A 3-level U-Net (16/32/64 channels, 191,841 FP32 master parameters, BF16 weights and
 * activations, FP32 accumulation on MFMA) that maps a rho grid deposited from few
 * macroparticles to the grid many more particles would deposit.  Images are FP32
 * [B][H][W] normalized to mean 1; H and W multiples of 64.  Trained with Adam by the
 * library's own forward and backward kernels. */
typedef struct ob_denoiser ob_denoiser;
int ob_dn_create(int H, int W, int max_batch, ob_denoiser** out);
int ob_dn_destroy(ob_denoiser* dn);
int ob_dn_num_params(ob_denoiser* dn, int64_t* n);
/* Parameters as one FP32 vector: per layer W[9][Cin][Cout] then b[Cout] (layers c1a c1b
 * c2a c2b c3a c3b c3c c3d c4a c4b c5a c5b), then the 1x1 head's 16 weights and bias.
 * Setting them resets the optimizer. */
int ob_dn_set_params(ob_denoiser* dn, const float* host, int64_t n);
int ob_dn_get_params(ob_denoiser* dn, float* host, int64_t n);
/* Inference on B images: host buffers, or device buffers (y_dev may equal x_dev). */
int ob_dn_forward(ob_denoiser* dn, const float* x, float* y, int B);
int ob_dn_forward_dev(ob_denoiser* dn, const float* x_dev, float* y_dev, int B);
/* MSE loss and the gradient of every parameter for B host images and targets. */
int ob_dn_loss_grad(ob_denoiser* dn, const float* x, const float* target, int B, float* loss,
                    float* grad);
/* One Adam step (lr) on B device-resident pairs; ob_dn_last_loss reads its loss. */
int ob_dn_train_step_dev(ob_denoiser* dn, const float* x_dev, const float* target_dev, int B,
                         float lr);
int ob_dn_last_loss(ob_denoiser* dn, float* loss);
/* Loss = mean (out - y)^2 + sum_{s=1..nscales} lam[s-1] * mean of the squared 2^s x 2^s
 * block means of out - y (low spatial frequencies carry the field error).  0: plain MSE. */
int ob_dn_set_loss(ob_denoiser* dn, int nscales, const float* lam);
/* Training pairs on the device: specs holds 5 floats per pair (family 0 Gaussian, 1 KV,
 * 2 waterbag, 3 hollow shell; x and y scales; cos and sin of the rotation).  x_dev gets the
 * n_low-particle TSC deposit, y_dev the n_high-particle one on the same grid (the square
 * box around the low-N sample), both normalized to mean 1. */
int ob_dn_make_pairs(int H, int W, int npairs, const float* specs, uint64_t seed,
                     int64_t n_low, int64_t n_high, float* x_dev, float* y_dev);
/* One training pair from a tracked bunch: x_dev gets the deposit of particles 0, stride,
 * 2 stride, ... and y_dev that of all particles, on the square box around the sample. */
int ob_dn_pair_from_bunch(int H, int W, ob_bunch* bunch, int64_t stride, float* x_dev,
                          float* y_dev);
/* dst image j <- src image idx[j] (< nsrc) for j < count (<= 64); images of n floats. */
int ob_dn_gather(const float* src, int64_t nsrc, float* dst, int64_t n, const int* idx,
                 int count);
/* Device float buffers and copies (kind 0 host->device, 1 device->host, 2 device->device). */
int ob_dn_alloc(int64_t nfloats, float** out);
int ob_dn_free(float* p);
int ob_dn_copy(float* dst, const float* src, int64_t nfloats, int kind);

/* SpaceChargeCalc2p5Drb with a denoiser: rho is replaced by the denoiser's estimate (same
 * total charge) after each deposit, before the moments and the Poisson solve.  The image
 * size must equal the grid; nullptr switches it off.  No reference counterpart. */
int ob_sc2p5drb_set_denoiser(ob_sc2p5drb* calc, ob_denoiser* dn);

/* ---- kernel timing (HIP events on the library stream) ---- */
/* Start (on != 0) or stop collecting per-kernel timings; clears earlier records. */
int ob_profile_enable(int on);
/* Lines "name count total_ms min_ms max_ms" for the collected launches. */
int ob_profile_report(char* buf, int buflen);
/* A plain streaming copy of n doubles (8n bytes read, 8n written) timed over reps launches
 * (no reference counterpart): the yardstick for the hardware byte counters. */
int ob_calibrate_copy(int64_t n, int reps, double* ms_per_copy);
/* The best streaming copy this library reaches (no reference counterpart): 16-byte loads,
 * 4 or 8 in flight per thread, plain or non-temporal stores, grids of 1024-8192
 * workgroups; ms per copy of n doubles (n even) and the variant index
 * 4 * grid + 2 * non-temporal + (8 in flight). */
int ob_calibrate_copy_wide(int64_t n, int reps, double* ms_per_copy, int* variant);

#ifdef __cplusplus
}
#endif
#endif
