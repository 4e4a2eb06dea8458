"""spacecharge -- space-charge calculators on the device.

Mirrors the reference module's SpaceChargeCalc2p5Drb
(src/spacecharge/wrap_spacechargecalc2p5d_rb.cc:38-51,189-198) and SpaceChargeCalc3D
(wrap_spacechargecalc3d.cc:38-45,171-177).  The grids returned by getRhoGrid() etc. are
host snapshots of the device grids taken at the call (the reference returns live views
of its host arrays); their accessors follow Grid1D/Grid2D/Grid3D (wrap_grid*.cc).
"""
import ctypes as C

import numpy as np

import _obnative as N


class Grid1D:
    """Grid1D(nZ[, length]) or Grid1D(nZ, zMin, zMax): a 1D cell-centred grid held on the
    host (src/spacecharge/Grid1D.cc, wrap_grid1D.cc:43-65).  Grid1D(nZ) spans [-1, 1] and
    Grid1D(nZ, length) spans [0, length], as the reference's wrapper builds them.
    binBunch / binBunchByParticle bin the bunch on the device (ob_grid1d_bin_bunch) and add
    the result to the grid; the calculators also return their line densities as Grid1D."""

    def __init__(self, nZ, *lims):
        nZ = int(nZ)
        if nZ < 1:
            raise N.OrbitError("Grid1D: the number of bins should be at least 1")
        if len(lims) == 0:
            zmin, zmax = -1.0, 1.0
        elif len(lims) == 1:
            zmin, zmax = 0.0, float(lims[0])
        elif len(lims) == 2:
            zmin, zmax = float(lims[0]), float(lims[1])
        else:
            raise TypeError("Grid1D(nZ[, length]) or Grid1D(nZ, zMin, zMax)")
        self._a = np.zeros(nZ)
        self._zmin, self._zmax = zmin, zmax

    @classmethod
    def _of(cls, arr, zMin, zMax):
        g = cls.__new__(cls)
        g._a = np.array(arr, dtype=np.float64)
        g._zmin, g._zmax = float(zMin), float(zMax)
        return g

    def setGridZ(self, zMin, zMax):
        self._zmin, self._zmax = float(zMin), float(zMax)

    def getSizeZ(self):
        return len(self._a)

    def getMinZ(self):
        return self._zmin

    def getMaxZ(self):
        return self._zmax

    def getStepZ(self):
        return (self._zmax - self._zmin) / len(self._a)

    def getGridZ(self, i):
        if i < 0 or i >= len(self._a):
            raise N.OrbitError("Grid1D.getGridZ: index out of range")
        return self._zmin + (i + 0.5) * self.getStepZ()

    def getValueOnGrid(self, i):
        if i < 0 or i >= len(self._a):
            raise N.OrbitError("Grid1D.getValueOnGrid: index out of range")
        return float(self._a[i])

    def setValue(self, value, i):
        if i < 0 or i >= len(self._a):
            raise N.OrbitError("Grid1D.setValue: index out of range")
        self._a[i] = value

    def setZero(self):
        self._a[:] = 0.0

    def multiply(self, coeff):
        self._a *= coeff

    def getSum(self):
        s = 0.0
        for v in self._a:  # Grid1D.cc:139-146, in index order
            s += float(v)
        return s

    def isInside(self, z):
        """1 inside [zMin, zMax], else 0 (Grid1D.cc:150-155)."""
        return 0 if (z < self._zmin or z > self._zmax) else 1

    def _ind_wz(self, z):
        # Grid1D::getIndAndWZ (Grid1D.cc:568-626): periodic at both ends
        n, dz = len(self._a), self.getStepZ()
        if n == 1:
            return 0, 0, 1.0, 0.0
        i0 = int(((z - self._zmin) / dz) - 0.5)
        frac = (z - (self._zmin + (i0 + 0.5) * dz)) / dz
        if i0 == 0 and frac < 0.0:
            return n - 1, 0, -frac, 1.0 + frac
        if i0 == n - 1:
            if frac > 0.0:
                return i0, 0, 1.0 - frac, frac
            return n - 2, i0, -frac, 1.0 + frac
        return i0, i0 + 1, 1.0 - frac, frac

    def binValue(self, value, z):
        """Grid1D::binValue (Grid1D.cc:422-440)."""
        if z < self._zmin or z > self._zmax:
            return
        i0, ip, w0, wp = self._ind_wz(z)
        if len(self._a) > 1:
            self._a[i0] += w0 * value
            self._a[ip] += wp * value
        else:
            self._a[i0] += value

    def getValue(self, z):
        """Grid1D::getValue (Grid1D.cc:191-210): 0 outside [zMin, zMax]."""
        if z < self._zmin or z > self._zmax:
            return 0.0
        i0, ip, w0, wp = self._ind_wz(z)
        if len(self._a) > 1:
            return float(w0 * self._a[i0] + wp * self._a[ip])
        return float(self._a[0])

    def calcGradient(self, z):
        """Grid1D::calcGradient (Grid1D.cc:520-534): (arr[ip] - arr[i0]) / dz between the
        two nodes z is binned to; 0 outside [zMin, zMax] and for a one-cell grid."""
        if z < self._zmin or z > self._zmax:
            return 0.0
        if len(self._a) == 1:
            return 0.0
        i0, ip, _, _ = self._ind_wz(z)
        return float((self._a[ip] - self._a[i0]) / self.getStepZ())

    def _bin(self, bunch, axis, by_particle):
        if not hasattr(bunch, "_h"):
            raise TypeError("Grid1D.binBunch: the first parameter must be a Bunch")
        bunch.compress()
        bunch._push_attrs()
        a = np.ascontiguousarray(self._a)
        N.call("ob_grid1d_bin_bunch", bunch._h, int(axis), int(by_particle), len(a),
               self._zmin, self._zmax, N.dptr(a))
        self._a = a

    def binBunch(self, bunch, axis=4):
        """binBunch(bunch[, axis]) (Grid1D.cc:241-275): the macrosize of each particle
        (attribute, else the bunch's) along coordinate axis (0..5 = x, xp, y, yp, z, dE),
        added to the grid."""
        self._bin(bunch, axis, False)

    def binBunchByParticle(self, bunch, axis=4):
        """binBunchByParticle(bunch[, axis]) (Grid1D.cc:335-352): weight 1 per particle."""
        self._bin(bunch, axis, True)

    def synchronizeMPI(self, comm=None):
        """One process holds the whole bunch: nothing to add across ranks."""
        return None

    def array(self):
        return self._a.copy()


class Grid2D:
    """2D node-centred grid, arr[ix][iy] (src/spacecharge/Grid2D.cc).

    Grid2D(xSize, ySize[, xMin, xMax, yMin, yMax]) makes a zero grid as the reference's
    constructor does (wrap_grid2D.cc; limits default to [-1, 1]); Grid2D(arr, xMin, xMax,
    yMin, yMax) wraps an array (the snapshots the calculators return)."""

    def __init__(self, arr, xMin=-1.0, xMax=1.0, yMin=-1.0, yMax=1.0, *rest):
        if isinstance(arr, (int, np.integer)):
            ySize = xMin
            lims = (xMax, yMin, yMax) + rest if rest else None
            self._a = np.zeros((int(arr), int(ySize)))
            self._l = tuple(float(v) for v in lims) if lims and len(lims) == 4 else (-1.0, 1.0, -1.0, 1.0)
            return
        self._a = np.array(arr, dtype=np.float64)
        self._l = (float(xMin), float(xMax), float(yMin), float(yMax))

    def setGridX(self, xMin, xMax):
        self._l = (float(xMin), float(xMax)) + self._l[2:]

    def setGridY(self, yMin, yMax):
        self._l = self._l[:2] + (float(yMin), float(yMax))

    def setValue(self, value, ix, iy):
        self._a[ix, iy] = value

    def setZero(self):
        self._a[:] = 0.0

    def getSizeX(self):
        return self._a.shape[0]

    def getSizeY(self):
        return self._a.shape[1]

    def getMinX(self):
        return self._l[0]

    def getMaxX(self):
        return self._l[1]

    def getMinY(self):
        return self._l[2]

    def getMaxY(self):
        return self._l[3]

    def getStepX(self):
        return (self._l[1] - self._l[0]) / (self.getSizeX() - 1)

    def getStepY(self):
        return (self._l[3] - self._l[2]) / (self.getSizeY() - 1)

    def getGridX(self, i):
        return self._l[0] + i * self.getStepX()

    def getGridY(self, i):
        return self._l[2] + i * self.getStepY()

    def getValueOnGrid(self, ix, iy):
        return float(self._a[ix, iy])

    def getSum(self):
        return float(self._a.sum())

    def array(self):
        return self._a.copy()


class Grid3D:
    """Snapshot of a 3D grid, arr[iz][ix][iy] (src/spacecharge/Grid3D.cc); z cell-centred."""

    def __init__(self, arr, lims):
        self._a = np.array(arr, dtype=np.float64)
        self._l = tuple(float(v) for v in lims)

    def getSizeX(self):
        return self._a.shape[1]

    def getSizeY(self):
        return self._a.shape[2]

    def getSizeZ(self):
        return self._a.shape[0]

    def getMinX(self):
        return self._l[0]

    def getMaxX(self):
        return self._l[1]

    def getMinY(self):
        return self._l[2]

    def getMaxY(self):
        return self._l[3]

    def getMinZ(self):
        return self._l[4]

    def getMaxZ(self):
        return self._l[5]

    def getStepX(self):
        return (self._l[1] - self._l[0]) / (self.getSizeX() - 1)

    def getStepY(self):
        return (self._l[3] - self._l[2]) / (self.getSizeY() - 1)

    def getStepZ(self):
        return (self._l[5] - self._l[4]) / self.getSizeZ()

    def getValueOnGrid(self, ix, iy, iz):
        return float(self._a[iz, ix, iy])

    def getSum(self):
        return float(self._a.sum())

    def array(self):
        return self._a.copy()


class PoissonSolverFFT2D:
    """PoissonSolverFFT2D(xSize, ySize[, xMin, xMax, yMin, yMax])
    (wrap_poissonsolverfft2d.cc:37-47): free-space 2D Poisson solver on the device, FP64
    Hockney convolution (PoissonSolverFFT2D.cc)."""

    def __init__(self, xSize, ySize, xMin=-1.0, xMax=1.0, yMin=-1.0, yMax=1.0):
        self._n = (int(xSize), int(ySize))
        self._l = (float(xMin), float(xMax), float(yMin), float(yMax))
        ratio = (self._l[1] - self._l[0]) / (self._l[3] - self._l[2])
        h = C.c_void_p()
        N.call("ob_sc2p5drb_create", self._n[0], self._n[1], 1, ratio, C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N._lib is not None:
            N._lib.ob_sc2p5drb_destroy(h)
            self._h = None

    def getSizeX(self):
        return self._n[0]

    def getSizeY(self):
        return self._n[1]

    def getMinX(self):
        return self._l[0]

    def getMaxX(self):
        return self._l[1]

    def getMinY(self):
        return self._l[2]

    def getMaxY(self):
        return self._l[3]

    def getStepX(self):
        return (self._l[1] - self._l[0]) / (self._n[0] - 1)

    def getStepY(self):
        return (self._l[3] - self._l[2]) / (self._n[1] - 1)

    def findPotential(self, rhoGrid2D, phiGrid2D):
        """phiGrid2D <- potential of rhoGrid2D (PoissonSolverFFT2D.cc:174-256).  The grids
        must have the solver's size and step ratio, and phi the rho grid's limits."""
        if (rhoGrid2D.getSizeX(), rhoGrid2D.getSizeY()) != self._n or \
                (phiGrid2D.getSizeX(), phiGrid2D.getSizeY()) != self._n or \
                phiGrid2D._l != rhoGrid2D._l:
            raise N.OrbitError("PoissonSolverFFT2D: the grid sizes or shape are different")
        rho = np.ascontiguousarray(rhoGrid2D._a, dtype=np.float64)
        phi = np.empty_like(rho)
        N.call("ob_sc2p5drb_find_potential", self._h, N.dptr(rho), *rhoGrid2D._l, N.dptr(phi))
        phiGrid2D._a[:] = phi


class SpaceChargeCalc2p5Drb:
    """SpaceChargeCalc2p5Drb(nx, ny, nz[, xy_ratio]) (SpaceChargeCalc2p5Drb.cc:31-63)."""

    def __init__(self, xSize, ySize, zSize, xy_ratio=1.0):
        self._n = (int(xSize), int(ySize), int(zSize))
        h = C.c_void_p()
        N.call("ob_sc2p5drb_create", self._n[0], self._n[1], self._n[2], float(xy_ratio),
               C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N._lib is not None:
            N._lib.ob_sc2p5drb_destroy(h)
            self._h = None

    def setDenoiser(self, denoiser):
        """Replace rho by the learned denoiser's estimate after every deposit (None: off).
        No reference counterpart; the denoiser's image size must equal the grid."""
        N.call("ob_sc2p5drb_set_denoiser", self._h, denoiser._h if denoiser is not None else None)
        self._denoiser = denoiser  # kept alive while in use

    def getDenoiser(self):
        return getattr(self, "_denoiser", None)

    def trackBunch(self, bunch, length, pipe_radius):
        """trackBunch(bunch, length, pipe_radius) (:110-168)."""
        bunch._push_attrs()
        N.call("ob_sc2p5drb_track", self._h, bunch._h, float(length), float(pipe_radius))

    def _grid(self, which):
        nx, ny, nz = self._n
        n = nx * ny if which <= 1 else nz
        out = np.zeros(n)
        lims = np.zeros(4)
        N.call("ob_sc2p5drb_get_grid", self._h, which, N.dptr(out), N.dptr(lims))
        return out, lims

    def getRhoGrid(self):
        a, l = self._grid(0)
        return Grid2D(a.reshape(self._n[0], self._n[1]), *l)

    def getPhiGrid(self):
        a, l = self._grid(1)
        return Grid2D(a.reshape(self._n[0], self._n[1]), *l)

    def getLongGrid(self):
        a, l = self._grid(2)
        return Grid1D._of(a, l[0], l[1])

    def getLongDerivativeGrid(self):
        a, l = self._grid(3)
        return Grid1D._of(a, l[0], l[1])

    def setLongAveragingPointsN(self, n):
        N.call("ob_sc2p5drb_set_long_avg_n", self._h, int(n))

    def getLongAveragingPointsN(self):
        v = C.c_int()
        N.call("ob_sc2p5drb_get_long_avg_n", self._h, C.byref(v))
        return int(v.value)

    def getMoments(self):
        """(x_center, y_center, a_bunch, totalMacrosize) of the last trackBunch."""
        out = np.zeros(4)
        N.call("ob_sc2p5drb_get_moments", self._h, N.dptr(out))
        return tuple(float(v) for v in out)


class Boundary2D:
    """Boundary2D(nPoints, nModes) or Boundary2D(nPoints, nModes, shape, xDim[, yDim])
    with shape "Circle", "Ellipse" or "Rectangle" (wrap_boundary2d.cc:38-75;
    BaseBoundary2D.cc, ShapedBoundary2D.cc)."""

    _SHAPES = ("no shape", "Circle", "Ellipse", "Rectangle")

    def __init__(self, nPoints, nModes, shape=None, xDim=None, yDim=None):
        h = C.c_void_p()
        if shape is None:
            N.call("ob_boundary2d_create", int(nPoints), int(nModes), C.byref(h))
        else:
            if xDim is None:
                raise TypeError("Boundary2D(nPoints, nModes, shape, xDim[, yDim])")
            yDim = xDim if yDim is None else yDim
            N.call("ob_boundary2d_create_shaped", int(nPoints), int(nModes),
                   str(shape).encode(), float(xDim), float(yDim), C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N._lib is not None:
            N._lib.ob_boundary2d_destroy(h)
            self._h = None

    def _info(self):
        lims = np.zeros(4)
        info = (C.c_int * 4)()
        N.call("ob_boundary2d_info", self._h, N.dptr(lims), info)
        return lims, list(info)

    def getMinX(self):
        return float(self._info()[0][0])

    def getMaxX(self):
        return float(self._info()[0][1])

    def getMinY(self):
        return float(self._info()[0][2])

    def getMaxY(self):
        return float(self._info()[0][3])

    def getNumberOfPoints(self):
        return self._info()[1][0]

    def getNumberOfModes(self):
        return self._info()[1][1]

    def isInitialized(self):
        return self._info()[1][2]

    def getShapeName(self):
        return self._SHAPES[self._info()[1][3]]

    def initialize(self):
        N.call("ob_boundary2d_initialize", self._h)

    def getBoundaryPoint(self, i):
        x, y = C.c_double(), C.c_double()
        N.call("ob_boundary2d_get_point", self._h, int(i), C.byref(x), C.byref(y))
        return (x.value, y.value)

    def setBoundaryPoint(self, i, x, y):
        N.call("ob_boundary2d_set_point", self._h, int(i), float(x), float(y))

    def isInside(self, x, y):
        v = C.c_int()
        N.call("ob_boundary2d_is_inside", self._h, float(x), float(y), C.byref(v))
        return v.value

    def addBoundaryPotential(self, rhoGrid, phiGrid):
        """Subtract the boundary's field from phiGrid in place (BaseBoundary2D.cc:349-458);
        rhoGrid must span the same grid."""
        if (rhoGrid.getSizeX(), rhoGrid.getSizeY(), rhoGrid.getMinX(), rhoGrid.getMinY(),
                rhoGrid.getStepX(), rhoGrid.getStepY()) != (
                phiGrid.getSizeX(), phiGrid.getSizeY(), phiGrid.getMinX(), phiGrid.getMinY(),
                phiGrid.getStepX(), phiGrid.getStepY()):
            raise N.OrbitError("Boundary2D addBoundaryPotential: the rho and phi grids differ")
        a = np.ascontiguousarray(phiGrid._a, dtype=np.float64)
        lims = np.array(phiGrid._l, dtype=np.float64)
        N.call("ob_boundary2d_add_potential", self._h, N.dptr(a), a.shape[0], a.shape[1],
               N.dptr(lims))
        phiGrid._a = a


class SpaceChargeCalc2p5D:
    """SpaceChargeCalc2p5D(nx, ny, nz[, xy_ratio]) (SpaceChargeCalc2p5D.cc:25-43,
    wrap_spacechargecalc2p5d.cc:41).  trackBunch(bunch, length) without a boundary."""

    def __init__(self, xSize, ySize, zSize, xy_ratio=1.0):
        self._n = (int(xSize), int(ySize), int(zSize))
        h = C.c_void_p()
        N.call("ob_sc2p5d_create", self._n[0], self._n[1], self._n[2], float(xy_ratio),
               C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N._lib is not None:
            N._lib.ob_sc2p5d_destroy(h)
            self._h = None

    def trackBunch(self, bunch, length, boundary=None):
        """trackBunch(bunch, length[, boundary]) (SpaceChargeCalc2p5D.cc:77-120,
        wrap_spacechargecalc2p5d.cc:128-150)."""
        bunch._push_attrs()
        if boundary is None:
            N.call("ob_sc2p5d_track", self._h, bunch._h, float(length))
        else:
            if not isinstance(boundary, Boundary2D):
                raise TypeError("SpaceChargeCalc2p5D.trackBunch: boundary must be a Boundary2D")
            N.call("ob_sc2p5d_track_boundary", self._h, bunch._h, float(length), boundary._h)

    def _grid(self, which):
        nx, ny, nz = self._n
        out = np.zeros(nx * ny if which <= 1 else nz)
        lims = np.zeros(4)
        N.call("ob_sc2p5d_get_grid", self._h, which, N.dptr(out), N.dptr(lims))
        return out, lims

    def getRhoGrid(self):
        a, l = self._grid(0)
        return Grid2D(a.reshape(self._n[0], self._n[1]), *l)

    def getPhiGrid(self):
        a, l = self._grid(1)
        return Grid2D(a.reshape(self._n[0], self._n[1]), *l)

    def getLongGrid(self):
        a, l = self._grid(2)
        return Grid1D._of(a, l[0], l[1])

    def getState(self):
        """(xy_ratio, solver box (xMin, xMax, yMin, yMax), totalMacrosize of the last step,
        Green function rebuilds): the calculator's private state, for diagnostics."""
        o = np.zeros(7)
        N.call("ob_sc2p5d_get_state", self._h, N.dptr(o))
        return float(o[0]), tuple(float(v) for v in o[1:5]), float(o[5]), int(o[6])


class SpaceChargeCalcSliceBySlice2D:
    """SpaceChargeCalcSliceBySlice2D(nx, ny, nz[, xy_ratio])
    (SpaceChargeCalcSliceBySlice2D.cc:15-32, wrap_spacechargecalc_slicebyslice_2D.cc:41)."""

    def __init__(self, xSize, ySize, zSize, xy_ratio=1.0):
        self._n = (int(xSize), int(ySize), int(zSize))
        h = C.c_void_p()
        N.call("ob_slice2d_create", self._n[0], self._n[1], self._n[2], float(xy_ratio),
               C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N._lib is not None:
            N._lib.ob_slice2d_destroy(h)
            self._h = None

    def trackBunch(self, bunch, length, boundary=None):
        """trackBunch(bunch, length[, boundary]) (:66-113)."""
        if boundary is not None and not isinstance(boundary, Boundary2D):
            raise TypeError("SpaceChargeCalcSliceBySlice2D.trackBunch: boundary must be a "
                            "Boundary2D")
        bunch._push_attrs()
        N.call("ob_slice2d_track", self._h, bunch._h, float(length),
               boundary._h if boundary is not None else None)

    def _grid(self, which):
        nx, ny, nz = self._n
        out = np.zeros(nx * ny * nz)
        lims = np.zeros(6)
        N.call("ob_slice2d_get_grid", self._h, which, N.dptr(out), N.dptr(lims))
        return Grid3D(out.reshape(nz, nx, ny), lims)

    def getRhoGrid(self):
        return self._grid(0)

    def getPhiGrid(self):
        return self._grid(1)

    def getState(self):
        """(xy_ratio, solver box (xMin, xMax, yMin, yMax), Green function rebuilds)."""
        o = np.zeros(6)
        N.call("ob_slice2d_get_state", self._h, N.dptr(o))
        return float(o[0]), tuple(float(v) for v in o[1:5]), int(o[5])


class SpaceChargeCalc3D:
    """SpaceChargeCalc3D(nx, ny, nz) (SpaceChargeCalc3D.cc:32-58)."""

    def __init__(self, xSize, ySize, zSize):
        self._n = (int(xSize), int(ySize), int(zSize))
        h = C.c_void_p()
        N.call("ob_sc3d_create", self._n[0], self._n[1], self._n[2], C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N._lib is not None:
            N._lib.ob_sc3d_destroy(h)
            self._h = None

    def trackBunch(self, bunch, length):
        """trackBunch(bunch, length) (:100-148)."""
        bunch._push_attrs()
        N.call("ob_sc3d_track", self._h, bunch._h, float(length))

    def _grid(self, which):
        nx, ny, nz = self._n
        out = np.zeros(nx * ny * nz)
        lims = np.zeros(6)
        N.call("ob_sc3d_get_grid", self._h, which, N.dptr(out), N.dptr(lims))
        return Grid3D(out.reshape(nz, nx, ny), lims)

    def getRhoGrid(self):
        return self._grid(0)

    def getPhiGrid(self):
        return self._grid(1)

    def setRatioLimit(self, r):
        N.call("ob_sc3d_set_ratio_limit", self._h, float(r))

    def getRatioLimit(self):
        v = C.c_double()
        N.call("ob_sc3d_get_ratio_limit", self._h, C.byref(v))
        return float(v.value)

    def numExtBunches(self, n=None):
        """numExtBunches([n]) (wrap_spacechargecalc3d.cc): n > 0 accounts for n neighbour
        bunches (wrappedBunchAnalysis, SpaceChargeCalc3D.cc:251-328); odd n rounds up."""
        if n is not None:
            N.call("ob_sc3d_set_num_ext_bunches", self._h, int(n))
        v = C.c_int()
        N.call("ob_sc3d_get_num_ext_bunches", self._h, C.byref(v))
        return int(v.value)

    def freqOfBunches(self, f=None):
        """freqOfBunches([f]): bunch frequency in Hz (default 402.5e6, :58)."""
        if f is not None:
            N.call("ob_sc3d_set_freq", self._h, float(f))
        v = C.c_double()
        N.call("ob_sc3d_get_freq", self._h, C.byref(v))
        return float(v.value)

    def setDepositMode(self, tiled=True, lds_budget_bytes=0):
        """Tile-sorted LDS deposit (default) or particle-order global atomics."""
        N.call("ob_sc3d_set_deposit_mode", self._h, 1 if tiled else 0, int(lds_budget_bytes))

    def getGreenRebuilds(self):
        v = C.c_int()
        N.call("ob_sc3d_green_rebuilds", self._h, C.byref(v))
        return int(v.value)
