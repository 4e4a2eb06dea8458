"""bunch -- the macroparticle container, device-resident on the MI355X.

Mirrors the reference's `bunch` extension module (src/orbit/wrap_bunch.cc:1178-1232,
src/orbit/wrap_syncpart.cc:549-566, BunchTwissAnalysis): same class and method names,
argument meaning and units (GeV, m, rad).  Coordinates live in HBM as six FP64 arrays;
per-particle accessors go through a host mirror and are correct but slow, so bulk code
should use setCoordinates()/getCoordinates().
"""
import sys

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

    def momentum(self, value=None):
        """momentum() -> GeV/c; momentum(p) sets p along z (wrap_syncpart.cc:136-156)."""
        if value is not None:
            self.momentum_set(value)
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

    def momentum_set(self, p):
        N.call("ob_bunch_set_sync_momentum", self._b._h, float(p))

    def rVector(self, v=None):
        """Position of the synchronous particle (host-side bookkeeping only)."""
        if v is not None:
            self._r = tuple(float(a) for a in v)
        return getattr(self, "_r", (0.0, 0.0, 0.0))

    def nxVector(self, v=None):
        """x-axis ort of the synchronous particle (host-side bookkeeping only)."""
        if v is not None:
            self._nx = tuple(float(a) for a in v)
        return getattr(self, "_nx", (1.0, 0.0, 0.0))

    def pVector(self, v=None):
        """Momentum vector; the tracked path uses its length (SyncPart.cc:152-164, 236-317):
        setting it sets momentum = |p|, and the vector is returned along z."""
        if v is not None:
            px, py, pz = (float(a) for a in v)
            self.momentum_set(float(np.sqrt(px * px + py * py + pz * pz)))
        return (0.0, 0.0, self.momentum())

    def x(self):
        return self.rVector()[0]

    def y(self):
        return self.rVector()[1]

    def z(self):
        return self.rVector()[2]

    def px(self):
        return self.pVector()[0]

    def py(self):
        return self.pVector()[1]

    def pz(self):
        return self.pVector()[2]

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

    def getMPIComm(self):
        """The bunch's communicator (Bunch.cc getMPI_Comm_Local): one process holds the whole
        bunch, so there is none."""
        return None

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

    def getMacroSizes(self):
        """Bulk read of the per-particle macrosize attribute (None without it)."""
        return None if self._msize is None else np.array(self._msize)

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
        """Append this bunch's particles to `other` (Bunch.cc:339-402).

        Both bunches must carry the same particle attributes (the reference aborts
        otherwise, :352-388).  Attribute values follow the reference exactly: the appended
        particles start at 0 (attrInit, :618-635, ParticleAttributes.cc:67-71), then the
        source's value of particle i is written to the target's particle i for i < the
        source's size (:390-401).  For an empty target that is a plain copy."""
        if self.getPartAttrNames() != other.getPartAttrNames():
            raise N.OrbitError("Bunch::addParticlesTo: the particles attributes have "
                               "different structures for two bunches")
        src_m = None if self._msize is None else np.array(self._msize)
        tgt_m = None if other._msize is None else np.array(other._msize)
        a = other.getCoordinates()
        other.setCoordinates(np.concatenate([a, self.getCoordinates()], axis=1))
        if src_m is not None:
            m = np.concatenate([tgt_m, np.zeros(len(src_m))])
            m[:len(src_m)] = src_m
            other.setMacroSizes(m)

    def flush(self):
        """Apply queued maps now (they are otherwise applied lazily)."""
        N.call("ob_bunch_flush", self._h)

    # ---- text files: dumpBunch / readBunch (Bunch.cc:894-1080, 177-284, 1116-1270;
    #      SyncPart.cc:323-540).  Same layout and number formatting as the reference:
    #      bunch attributes with 6 significant digits, the synchronous particle with 14,
    #      coordinates with 8 (C++ ostream default float format, i.e. printf %g). ------
    def dumpBunch(self, fileName=None):
        """dumpBunch() prints to stdout; dumpBunch(fileName) writes the file."""
        import sys
        text = self._dump_text()
        if fileName is None:
            sys.stdout.write(text)
            sys.stdout.flush()
        else:
            with open(fileName, "w") as f:
                f.write(text)

    def _dump_text(self):
        g6 = lambda v: "%.6g" % v
        g14 = lambda v: "%.14g" % v
        g8 = lambda v: "%.8g" % v
        has_m = self._msize is not None
        out = ["% PARTICLE_ATTRIBUTES_CONTROLLERS_NAMES " + ("macrosize " if has_m else "") + "\n"]
        for name in sorted(self._int_attrs):
            out.append("% BUNCH_ATTRIBUTE_INT " + name + "   " + "%d" % self._int_attrs[name] + " \n")
        dbl = dict(self._dbl_attrs)
        dbl.update({"mass": self.mass(), "charge": self.charge(),
                    "classical_radius": self.classicalRadius(), "macro_size": self.macroSize()})
        for name in sorted(dbl):
            out.append("% BUNCH_ATTRIBUTE_DOUBLE " + name + "   " + g6(dbl[name]) + " \n")
        sp = self._sync
        px, py, pz = sp.pVector()
        out.append("%  SYNC_PART_COORDS " + " ".join(g14(v) for v in sp.rVector()) +
                   "  x, y, z positions in [m]\n")
        out.append("%  SYNC_PART_MOMENTUM " + " ".join(g14(v) for v in (px, py, pz)) +
                   "  px, py, pz momentum component in GeV/c\n")
        out.append("%  SYNC_PART_X_AXIS " + " ".join(g14(v) for v in sp.nxVector()) +
                   "  nxx, nxy, pxz - x-axis ort coordinates\n")
        out.append("%  info only: energy of the synchronous particle [GeV] = " +
                   g14(sp.kinEnergy()) + " \n")
        out.append("%  info only: momentum of the synchronous particle [GeV/c] = " +
                   g14(sp.momentum()) + " \n")
        out.append("%  info only: beta=v/c of the synchronous particle = " + g14(sp.beta()) + " \n")
        out.append("%  info only: gamma=1/sqrt(1-(v/c)**2) of the synchronous particle = " +
                   g14(sp.gamma()) + " \n")
        out.append("%  SYNC_PART_TIME " + g14(sp.time()) + "  time in [sec]\n")
        out.append("% x[m] px[rad] y[m] py[rad] z[m]  (pz or dE [GeV]) " +
                   ("macro_size " if has_m else "") + "\n")
        c = self.getCoordinates()
        cols = [c[k] for k in range(6)] + ([np.asarray(self._msize)] if has_m else [])
        for row in zip(*cols):
            out.append("".join(g8(v) + " " for v in row) + "\n")
        return "".join(out)

    def readBunch(self, fileName, nParts=None):
        """readBunch(fileName[, nParts]): bunch attributes, synchronous particle, then the
        particles (appended; at most nParts).  A data line holds 6 values, or 6 plus one
        per particle attribute the bunch already has; anything else is an error, as in the
        reference (which aborts)."""
        with open(fileName) as f:
            lines = f.read().split("\n")
        # initBunchAttributes: the leading '%' lines (Bunch.cc:177-284)
        self._int_attrs = {}
        self._dbl_attrs = {}
        for ln in lines:
            if not ln or ln[0] != "%":
                break
            t = ln.split()
            if len(t) > 3 and t[1] == "BUNCH_ATTRIBUTE_INT":
                self._int_attrs[t[2]] = int(float(t[3]))
            elif len(t) > 3 and t[1] == "BUNCH_ATTRIBUTE_DOUBLE":
                v = float(t[3])
                setter = {"mass": self.mass, "charge": self.charge,
                          "classical_radius": self.classicalRadius,
                          "macro_size": self.macroSize}.get(t[2])
                if setter is not None:
                    setter(v)
                else:
                    self._dbl_attrs[t[2]] = v
        # readSyncPart (SyncPart.cc:323-460)
        sp = self._sync
        for ln in lines:
            if not ln or ln[0] != "%":
                break
            t = ln.split()
            if len(t) > 5 and t[1] == "SYNC_PART_COORDS":
                sp.rVector((float(t[2]), float(t[3]), float(t[4])))
            elif len(t) > 5 and t[1] == "SYNC_PART_MOMENTUM":
                sp.pVector((float(t[2]), float(t[3]), float(t[4])))
            elif len(t) > 5 and t[1] == "SYNC_PART_X_AXIS":
                sp.nxVector((float(t[2]), float(t[3]), float(t[4])))
            elif len(t) > 3 and t[1] == "SYNC_PART_TIME":
                sp.time(float(t[2]))
        # readBunchCoords (Bunch.cc:1116-1270)
        n_attr = 1 if self._msize is not None else 0
        rows = []
        for k, ln in enumerate(lines):
            if not ln or ln[0] == "%":
                continue
            t = ln.split()
            if len(t) != 6 + n_attr and len(t) != 6:
                raise N.OrbitError(f"Bunch.readBunch: {fileName}: error in data structure in "
                                   f"line {k + 1}: {ln!r}")
            rows.append([float(v) for v in t] + ([0.0] * n_attr if len(t) == 6 else []))
            if nParts is not None and len(rows) >= int(nParts):
                break
        if not rows:
            return self.getSize()
        data = np.array(rows, dtype=np.float64).T
        old = self.getCoordinates()
        m_old = None if self._msize is None else np.asarray(self._msize)
        self.setCoordinates(np.concatenate([old, data[:6]], axis=1))
        if n_attr:
            self.setMacroSizes(np.concatenate([m_old, data[6]]))
        return self.getSize()

    # ---- full-precision binary checkpoint (no reference counterpart) --------
    def saveBunchBinary(self, fileName):
        """Exact checkpoint: coordinates, macrosizes, bunch attributes and the synchronous
        particle, bit for bit (numpy .npz, nothing executable)."""
        sp = self._sync
        meta = np.array([self.mass(), self.charge(), self.classicalRadius(), self.macroSize(),
                         sp.momentum(), sp.time()], dtype=np.float64)
        arrays = {"coords": self.getCoordinates(), "meta": meta,
                  "sync_r": np.array(sp.rVector()), "sync_nx": np.array(sp.nxVector()),
                  "dbl_names": np.array(sorted(self._dbl_attrs), dtype=np.str_),
                  "dbl_vals": np.array([self._dbl_attrs[k] for k in sorted(self._dbl_attrs)]),
                  "int_names": np.array(sorted(self._int_attrs), dtype=np.str_),
                  "int_vals": np.array([self._int_attrs[k] for k in sorted(self._int_attrs)],
                                       dtype=np.int64)}
        if self._msize is not None:
            arrays["macrosizes"] = np.asarray(self._msize)
        with open(fileName, "wb") as f:
            np.savez(f, **arrays)

    def readBunchBinary(self, fileName):
        """Restore a checkpoint written by saveBunchBinary (replaces the particles)."""
        with np.load(fileName, allow_pickle=False) as z:
            m = z["meta"]
            self.mass(m[0])
            self.charge(m[1])
            self.classicalRadius(m[2])
            self.macroSize(m[3])
            self._sync.momentum(m[4])
            self._sync.time(m[5])
            self._sync.rVector(tuple(z["sync_r"]))
            self._sync.nxVector(tuple(z["sync_nx"]))
            self._dbl_attrs = {str(k): float(v) for k, v in zip(z["dbl_names"], z["dbl_vals"])}
            self._int_attrs = {str(k): int(v) for k, v in zip(z["int_names"], z["int_vals"])}
            self.setCoordinates(z["coords"])
            if "macrosizes" in z.files:
                self.setMacroSizes(z["macrosizes"])
            else:
                self.removePartAttr("macrosize")
        return self.getSize()


class BunchTwissAnalysis:
    """RMS moments, emittances and Twiss parameters of a bunch
    (src/orbit/BunchDiagnostics/BunchTwissAnalysis.cc:49-142, 200-420).

    analyzeBunch compresses the bunch (BunchTwissAnalysis.cc:63) and reads the 28 weighted
    sums of ob_bunch_moments, computed on the device after the queued maps: the
    coordinates are not downloaded.  The getters derive everything from those sums with
    the reference's formulas, including its choice of absolute values and the dispersive
    corrections of the x and y planes."""

    def __init__(self):
        self._avg = np.zeros(6)
        self._corr = np.zeros((6, 6))
        self._count = 0
        self._total = 0.0
        self._p = 0.0
        self._gb = (1.0, 1.0)

    def analyzeBunch(self, bunch):
        bunch.compress()
        bunch._push_attrs()
        s = np.zeros(28)
        N.call("ob_bunch_moments", bunch._h, N.dptr(s))
        self._count = bunch.getSize()
        self._total = float(s[0])
        raw = np.zeros((6, 6))
        k = 7
        for a in range(6):
            for b in range(a, 6):
                raw[b, a] = s[k]  # the reference fills corr[i + 6 j] for j <= i only
                k += 1
        if abs(self._total) > 0.0:
            self._avg = s[1:7] / self._total
            low = raw / self._total
            self._corr = low + np.tril(low, -1).T
        else:  # BunchTwissAnalysis.cc:117-128 leaves the unnormalized sums in place
            self._avg = s[1:7].copy()
            self._corr = raw
        sp = bunch.getSyncParticle()
        self._gb = (sp.gamma(), sp.beta())
        self._p = sp.momentum()
        self._ek_m = (sp.kinEnergy(), sp.mass())

    def computeBunchMoments(self, bunch, order, dispersionflag, emitnormflag):
        """computeBunchMoments(bunch, order, dispterm, emitnormterm)
        (BunchTwissAnalysis.cc:145-337, wrap_bunch_twiss_analysis.cc:56-75).

        Runs analyzeBunch, then sums on the device, for j < order and i <= order - j,
        (x - D_x dE / (E_kin + M) / beta^2 - <x>)^i (y - <y>)^j / (xb^i yb^j) per particle
        (the dispersion term only for dispersionflag > 0; xb, yb = sqrt(beta) or, for
        emitnormflag > 0, sqrt(beta emittance)).  The reference's own arithmetic is kept:
        with the macrosize attribute each factor carries the particle's macrosize and every
        term is added a second time without the xb, yb normalisation; without it entry
        [2][order] collects (x - <x>)^2 (y - <y>)^order undivided.  Entries with
        i + j <= order are divided by the total macrosize; [0][0] = 1, [1][0] = <x>,
        [0][1] = <y>, and (y - <y>)^order stays 0.  Orders 1 to 12 (order 0 makes the
        reference write outside its array)."""
        order = int(order)
        if order < 1 or order > 12:
            raise N.OrbitError("BunchTwissAnalysis.computeBunchMoments: order must be "
                               "between 1 and 12")
        self.analyzeBunch(bunch)
        x_avg, y_avg = self.getAverage(0), self.getAverage(2)
        with np.errstate(invalid="ignore"):
            if emitnormflag > 0:
                xb = float(np.sqrt(np.float64(self.getBeta(0) * self.getEmittance(0))))
                yb = float(np.sqrt(np.float64(self.getBeta(1) * self.getEmittance(1))))
            else:
                xb = float(np.sqrt(np.float64(self.getBeta(0))))
                yb = float(np.sqrt(np.float64(self.getBeta(1))))
        beta = self._gb[1]
        prm = np.array([x_avg, y_avg, self.getDispersion(0), self._ek_m[0] + self._ek_m[1],
                        beta * beta, xb, yb, 1.0 if dispersionflag > 0 else 0.0])
        out = np.zeros(order * (order + 3) // 2 + 2)
        N.call("ob_bunch_xy_moments", bunch._h, order, N.dptr(prm), N.dptr(out))
        mom = np.zeros((order + 1, order + 1))
        k = 0
        for j in range(order):
            for i in range(order + 1 - j):
                mom[i, j] = out[k]
                k += 1
        if order >= 2:
            mom[2, order] = out[k]
        total = out[k + 1]
        if abs(total) > 0.0:
            for i in range(order + 1):
                mom[i, :order + 1 - i] /= total
        mom[0, 0] = 1.0
        mom[1, 0] = x_avg
        mom[0, 1] = y_avg
        self._mom = mom

    def getBunchMoment(self, i, j):
        """momentXY[i][j] of the last computeBunchMoments (BunchTwissAnalysis.cc:350-353);
        0 outside 0..order."""
        mom = getattr(self, "_mom", None)
        if mom is None:
            return 0.0
        o = mom.shape[0] - 1
        if i < 0 or i > o or j < 0 or j > o:
            return 0.0
        return float(mom[i, j])

    def getGlobalCount(self):
        return self._count

    def getGlobalMacrosize(self):
        return self._total

    def getAverage(self, i):
        if i < 0 or i > 5:
            return 0.0
        return float(self._avg[i])

    def getCorrelation(self, i, j):
        if i < 0 or i > 5 or j < 0 or j > 5:
            return 0.0
        return float(self._corr[i, j] - self._avg[i] * self._avg[j])

    def _plane(self, ic):
        g = self.getCorrelation
        return (abs(g(2 * ic, 2 * ic)), abs(g(2 * ic + 1, 2 * ic + 1)), g(2 * ic, 2 * ic + 1),
                g(2 * ic, 5), g(2 * ic + 1, 5), abs(g(5, 5)))

    def getEmittance(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        x2, xp2, xxp, xdE, xpdE, dE2 = self._plane(ic)
        if ic == 2 or dE2 == 0:
            return float(np.sqrt(abs(x2 * xp2 - xxp * xxp)))
        a = x2 - xdE * xdE / dE2
        b = xp2 - xpdE * xpdE / dE2
        c = xxp - xdE * xpdE / dE2
        return float(np.sqrt(abs(a * b - c * c)))

    def getEmittanceNormalized(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        if ic == 2:
            return self.getEmittance(ic)
        return self.getEmittance(ic) * self._gb[0] * self._gb[1]

    # Twiss parameters without the dispersive part for x and y (BunchTwissAnalysis.cc:257-306).
    # The reference divides by <dE^2> and by the emittance unguarded; so does this (IEEE
    # division: inf or nan where the reference's C arithmetic gives the same).
    def getAlpha(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        x2, xp2, xxp, xdE, xpdE, dE2 = self._plane(ic)
        num = xxp if ic == 2 else xxp - _div(xdE * xpdE, dE2)
        return -_div(num, self.getEmittance(ic))

    def getBeta(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        x2, xp2, xxp, xdE, xpdE, dE2 = self._plane(ic)
        num = x2 if ic == 2 else x2 - _div(xdE * xdE, dE2)
        return _div(num, self.getEmittance(ic))

    def getGamma(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        x2, xp2, xxp, xdE, xpdE, dE2 = self._plane(ic)
        num = xp2 if ic == 2 else xp2 - _div(xpdE * xpdE, dE2)
        return _div(num, self.getEmittance(ic))

    def getDispersion(self, ic):
        if ic < 0 or ic > 1:
            return 0.0
        return _div(self.getCorrelation(2 * ic, 5), abs(self.getCorrelation(5, 5))) \
            * self._p * self._gb[1]

    def getDispersionDerivative(self, ic):
        if ic < 0 or ic > 1:
            return 0.0
        return _div(self.getCorrelation(2 * ic + 1, 5), abs(self.getCorrelation(5, 5))) \
            * self._p * self._gb[1]

    # Effective (dispersion-inclusive) values (BunchTwissAnalysis.cc:340-420).
    def _eff(self, ic):
        x2, xp2, xxp = self._plane(ic)[:3]
        return x2, xp2, xxp, x2 * xp2 - xxp * xxp

    def getEffectiveEmittance(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        return float(np.sqrt(abs(self._eff(ic)[3])))

    def getEffectiveAlpha(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        x2, xp2, xxp, e2 = self._eff(ic)
        return 0.0 if e2 <= 0.0 else -xxp / float(np.sqrt(e2))

    def getEffectiveBeta(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        x2, xp2, xxp, e2 = self._eff(ic)
        return 0.0 if e2 <= 0.0 else x2 / float(np.sqrt(e2))

    def getEffectiveGamma(self, ic):
        if ic < 0 or ic > 2:
            return 0.0
        x2, xp2, xxp, e2 = self._eff(ic)
        return sys.float_info.max if e2 <= 0.0 else xp2 / float(np.sqrt(e2))


def _div(a, b):
    """a / b with C's IEEE semantics (x/0 = +-inf, 0/0 = nan) instead of Python's error."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float64(a) / np.float64(b))
