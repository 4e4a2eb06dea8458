"""bunch -- the macroparticle container, device-resident on the MI355X.

Mirrors the reference's `bunch` extension module (src/orbit/wrap_bunch.cc:1178-1232,
src/orbit/wrap_syncpart.cc:549-566, BunchTwissAnalysis): same class and method names,
argument meaning and units (GeV, m, rad).  Coordinates live in HBM as six FP64 arrays;
per-particle accessors go through a host mirror and are correct but slow, so bulk code
should use setCoordinates()/getCoordinates().
"""
import numpy as np

import _obnative as N
import ctypes as C

_COORD = {"x": 0, "xp": 1, "px": 1, "y": 2, "yp": 3, "py": 3, "z": 4, "dE": 5, "pz": 5}


class SyncPart:
    """Synchronous particle (src/orbit/SyncPart.cc; Python methods wrap_syncpart.cc:549-566)."""

    def __init__(self, bunch):
        self._b = bunch

    def _get(self):
        out = np.zeros(6)
        N.call("ob_bunch_get_sync", self._b._h, N.dptr(out))
        return out

    def kinEnergy(self, value=None):
        """kinEnergy() -> GeV; kinEnergy(T) sets pz from T (wrap_syncpart.cc:70-92)."""
        if value is not None:
            N.call("ob_bunch_set_kin_energy", self._b._h, float(value))
        return float(self._get()[0])

    def momentum(self):
        return float(self._get()[1])

    def beta(self):
        return float(self._get()[2])

    def gamma(self):
        return float(self._get()[3])

    def time(self, value=None):
        if value is not None:
            N.call("ob_bunch_set_sync_time", self._b._h, float(value))
        return float(self._get()[4])

    def mass(self):
        return float(self._get()[5])

    def energyToMomentum(self, ek):  # SyncPart.cc energyToMomentum
        m = self.mass()
        return float(np.sqrt(ek * (ek + 2.0 * m)))

    def momentumToEnergy(self, p):  # SyncPart.cc momentumToEnergy
        m = self.mass()
        return float(p * p / (np.sqrt(m * m + p * p) + m))


class Bunch:
    """Bunch() -- protons, macroSize 0 (src/orbit/Bunch.cc:404-425)."""

    def __init__(self):
        h = C.c_void_p()
        N.call("ob_bunch_create", C.byref(h))
        self._h = h
        self._sync = SyncPart(self)
        self._msize = None          # per-particle "macrosize" attribute (host copy)
        self._msize_dirty = False
        self._int_attrs = {}
        self._dbl_attrs = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N._lib is not None:
            N._lib.ob_bunch_destroy(h)
            self._h = None

    # ---- bunch attributes -------------------------------------------------
    def _attrs(self):
        out = np.zeros(4)
        N.call("ob_bunch_get_attrs", self._h, N.dptr(out))
        return out

    def mass(self, value=None):
        if value is not None:
            N.call("ob_bunch_set_mass", self._h, float(value))
        return float(self._attrs()[0])

    def charge(self, value=None):
        if value is not None:
            N.call("ob_bunch_set_charge", self._h, float(value))
        return float(self._attrs()[1])

    def classicalRadius(self, value=None):
        if value is not None:
            N.call("ob_bunch_set_classical_radius", self._h, float(value))
        return float(self._attrs()[2])

    def macroSize(self, value=None):
        if value is not None:
            N.call("ob_bunch_set_macrosize", self._h, float(value))
        return float(self._attrs()[3])

    def getSyncParticle(self):
        return self._sync

    def bunchAttrInt(self, name, value=None):
        if value is not None:
            self._int_attrs[name] = int(value)
        return self._int_attrs[name]

    def hasBunchAttrInt(self, name):
        return 1 if name in self._int_attrs else 0

    def bunchAttrDouble(self, name, value=None):
        if value is not None:
            self._dbl_attrs[name] = float(value)
        return self._dbl_attrs[name]

    def hasBunchAttrDouble(self, name):
        return 1 if name in self._dbl_attrs else 0

    # ---- particles ----------------------------------------------------------
    def addParticle(self, x, xp, y, yp, z, dE):
        N.call("ob_bunch_add_particle", self._h, float(x), float(xp), float(y), float(yp),
               float(z), float(dE))
        if self._msize is not None:
            self._msize = np.append(self._msize, 0.0)
            self._msize_dirty = True

    def getSize(self):
        n = C.c_int64()
        N.call("ob_bunch_get_size", self._h, C.byref(n))
        return int(n.value)

    def getSizeGlobal(self):
        """One process holds the whole bunch: equals getSize() (Bunch.cc:850-861)."""
        return self.getSize()

    def getTotalCount(self):
        return self.getSize()

    def _coord(self, k, i, value):
        i = int(i)
        if value is not None:
            N.call("ob_bunch_set_coord", self._h, i, k, float(value))
            return float(value)
        v = C.c_double()
        N.call("ob_bunch_get_coord", self._h, i, k, C.byref(v))
        return float(v.value)

    def x(self, i, value=None):
        return self._coord(0, i, value)

    def xp(self, i, value=None):
        return self._coord(1, i, value)

    px = xp

    def y(self, i, value=None):
        return self._coord(2, i, value)

    def yp(self, i, value=None):
        return self._coord(3, i, value)

    py = yp

    def z(self, i, value=None):
        return self._coord(4, i, value)

    def dE(self, i, value=None):
        return self._coord(5, i, value)

    pz = dE

    def deleteParticle(self, i):
        N.call("ob_bunch_delete_particle", self._h, int(i))
        self.compress()

    def deleteParticleFast(self, i):
        N.call("ob_bunch_delete_particle", self._h, int(i))

    def deleteAllParticles(self):
        self.setCoordinates(np.zeros((6, 0)))

    def compress(self):
        """Remove deleted particles, keeping order (Bunch.cc:704-763)."""
        self._push_attrs()
        N.call("ob_bunch_compress", self._h)
        if self._msize is not None:  # the attribute is compressed natively with the particles
            m = np.zeros(self.getSize())
            N.call("ob_bunch_get_macrosizes", self._h, len(m), N.dptr(m))
            self._msize = m

    # ---- bulk SoA access (no reference counterpart: the fast path) ----------
    def setCoordinates(self, coords):
        """Replace all particles by a (6, n) array of x, xp, y, yp, z, dE."""
        a = np.ascontiguousarray(np.asarray(coords, dtype=np.float64))
        if a.ndim != 2 or a.shape[0] != 6:
            raise ValueError("coords must have shape (6, n)")
        n = a.shape[1]
        N.call("ob_bunch_set_coords", self._h, n, *(N.dptr(a[k]) for k in range(6)))
        if self._msize is not None:
            self._msize = np.zeros(n)
            self._msize_dirty = True

    def getCoordinates(self):
        """Return all coordinates as a new (6, n) float64 array."""
        n = self.getSize()
        out = np.zeros((6, n))
        N.call("ob_bunch_get_coords", self._h, *(N.dptr(out[k]) for k in range(6)))
        return out

    # ---- particle attributes: only "macrosize" is used on this path ---------
    def addPartAttr(self, name, params=None):
        if name != "macrosize":
            raise NotImplementedError(f"particle attribute '{name}' is not supported")
        if self._msize is None:
            self._msize = np.zeros(self.getSize())
            self._msize_dirty = True

    def removePartAttr(self, name):
        if name == "macrosize":
            self._msize = None
            N.call("ob_bunch_set_macrosizes", self._h, 0, C.cast(None, C.POINTER(C.c_double)))

    def removeAllPartAttr(self):
        self.removePartAttr("macrosize")

    def hasPartAttr(self, name):
        return 1 if (name == "macrosize" and self._msize is not None) else 0

    def getPartAttrNames(self):
        return ("macrosize",) if self._msize is not None else ()

    def partAttrValue(self, name, i, k, value=None):
        if name != "macrosize" or self._msize is None:
            raise KeyError(f"bunch has no particle attribute '{name}'")
        if k != 0:
            raise IndexError("macrosize has one value per particle")
        if value is not None:
            self._msize[int(i)] = float(value)
            self._msize_dirty = True
        return float(self._msize[int(i)])

    def setMacroSizes(self, m):
        """Bulk set of the per-particle macrosize attribute (adds it if missing)."""
        m = np.ascontiguousarray(np.asarray(m, dtype=np.float64))
        if m.shape != (self.getSize(),):
            raise ValueError("one macrosize per particle expected")
        self._msize = m.copy()
        self._msize_dirty = True
        self._push_attrs()

    def _push_attrs(self):
        if self._msize_dirty and self._msize is not None:
            N.call("ob_bunch_set_macrosizes", self._h, len(self._msize), N.dptr(self._msize))
        self._msize_dirty = False

    # ---- copies ------------------------------------------------------------
    def copyEmptyBunchTo(self, other):
        other.mass(self.mass())
        other.charge(self.charge())
        other.classicalRadius(self.classicalRadius())
        other.macroSize(self.macroSize())
        other.getSyncParticle().kinEnergy(self._sync.kinEnergy())
        other.getSyncParticle().time(self._sync.time())

    def copyBunchTo(self, other):
        self.copyEmptyBunchTo(other)
        other.setCoordinates(self.getCoordinates())
        if self._msize is not None:
            other.setMacroSizes(self._msize)

    def addParticlesTo(self, other):
        a = other.getCoordinates()
        other.setCoordinates(np.concatenate([a, self.getCoordinates()], axis=1))

    def flush(self):
        """Apply queued maps now (they are otherwise applied lazily)."""
        N.call("ob_bunch_flush", self._h)


class BunchTwissAnalysis:
    """RMS moments and emittances (src/orbit/BunchDiagnostics/BunchTwissAnalysis.cc:49-142,
    344-391).  Computed in FP64 on the downloaded coordinates."""

    def __init__(self):
        self._avg = np.zeros(6)
        self._corr = np.zeros((6, 6))
        self._count = 0
        self._total = 0.0
        self._gb = (1.0, 1.0)

    def analyzeBunch(self, bunch):
        c = bunch.getCoordinates()
        n = c.shape[1]
        if bunch.hasPartAttr("macrosize"):
            w = np.array(bunch._msize)
        else:
            w = np.ones(n)
        tot = float(w.sum())
        self._count, self._total = n, tot
        if tot != 0.0:
            self._avg = (c * w).sum(axis=1) / tot
            self._corr = (c * w) @ c.T / tot
        sp = bunch.getSyncParticle()
        self._gb = (sp.gamma(), sp.beta())

    def getGlobalCount(self):
        return self._count

    def getGlobalMacrosize(self):
        return self._total

    def getAverage(self, i):
        return float(self._avg[i])

    def getCorrelation(self, i, j):
        return float(self._corr[i, j] - self._avg[i] * self._avg[j])

    def getEmittance(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        g = self.getCorrelation
        x2, xp2, xxp = abs(g(2 * ic, 2 * ic)), abs(g(2 * ic + 1, 2 * ic + 1)), g(2 * ic, 2 * ic + 1)
        xdE, xpdE, dE2 = g(2 * ic, 5), g(2 * ic + 1, 5), abs(g(5, 5))
        if ic == 2 or dE2 == 0:
            return float(np.sqrt(abs(x2 * xp2 - xxp * xxp)))
        a = x2 - xdE * xdE / dE2
        b = xp2 - xpdE * xpdE / dE2
        c = xxp - xdE * xpdE / dE2
        return float(np.sqrt(abs(a * b - c * c)))

    def getEmittanceNormalized(self, ic):
        if ic == 2:
            return self.getEmittance(ic)
        return self.getEmittance(ic) * self._gb[0] * self._gb[1]
