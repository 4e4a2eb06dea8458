"""batch -- many independent bunches tracked together (no reference counterpart).

    from batch import BunchBatch
    bb = BunchBatch(bunches, calcs)          # one SpaceChargeCalc2p5Drb per bunch
    setSC2p5DrbAccNodes(lat, 0.49, bb.calculator(), pipe_radius)
    lat.trackBunch(bb.bunch())                # one lattice pass tracks every bunch
    bb.capture(lambda: lat.trackBunch(bb.bunch()))
    bb.replay(99)                             # the captured cell as a hipGraph, 99 times

All bunches share the lattice and the synchronous particle (same mass, charge, classical
radius and kinetic energy); their sizes, distributions, intensities and SC grids may
differ.  Each phase of a step is one launch over all bunches; the results are those of
tracking each bunch on its own.  While a bunch is in a batch its maps go through the batch
(bb.bunch() for the lattice, trackSequence) and reading it (getCoordinates, x(i), ...)
first applies the batch's pending work.
"""
import ctypes as C
import weakref

import _obnative as N
import teapot_base as TPB


class BunchBatch:
    def __init__(self, bunches, calcs=None):
        if not bunches:
            raise ValueError("a batch needs at least one bunch")
        for b in bunches:
            b._push_attrs()
        arr = (C.c_void_p * len(bunches))(*[b._h for b in bunches])
        h = C.c_void_p()
        N.call("ob_batch_create", arr, len(bunches), C.byref(h))
        self._h = h
        self.bunches = list(bunches)
        self.calcs = None
        self._proxy = _BatchBunch(self)
        self._calc = _BatchCalc(self)
        if calcs is not None:
            self.setCalculators(calcs)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._lib is not None:
            N._lib.ob_batch_destroy(h)
            self._h = None

    def setCalculators(self, calcs):
        if len(calcs) != len(self.bunches):
            raise ValueError("one calculator per bunch")
        arr = (C.c_void_p * len(calcs))(*[c._h for c in calcs])
        N.call("ob_batch_set_sc2p5drb", self._h, arr)
        self.calcs = list(calcs)

    def bunch(self):
        """Stand-in bunch for lattice tracking: its maps are applied to every bunch."""
        return self._proxy

    def calculator(self):
        """Stand-in SpaceChargeCalc2p5Drb for the SC nodes: one batched step per call."""
        return self._calc

    def trackSequence(self, ops):
        """Queue maps for every bunch, e.g. [("drift", 0.5), ("quad1", 0.25, 0.5)]."""
        keep = []
        arr = (TPB._MapOp * len(ops))(*[TPB._map_op(op[0], *op[1:], keep=keep) for op in ops])
        N.call("ob_batch_track_sequence", self._h, arr, len(ops))

    def flush(self):
        N.call("ob_batch_flush", self._h)

    def synchronize(self):
        """Wait; raises if a bunch had zero extent in its last SC step."""
        N.call("ob_batch_synchronize", self._h)

    def tuning(self):
        c, l, t = C.c_int(), C.c_int(), C.c_int()
        N.call("ob_batch_get_tuning", self._h, C.byref(c), C.byref(l), C.byref(t))
        return {"chunk": c.value, "deposit_item": l.value, "tuned": bool(t.value)}

    def setTuning(self, chunk, deposit_item):
        """Particles per push/hist work item and per tile-deposit work item."""
        N.call("ob_batch_set_tuning", self._h, int(chunk), int(deposit_item))

    def capture(self, fn):
        """Record the batch work done by fn() as a hipGraph (not run: replay runs it)."""
        N.call("ob_batch_capture_begin", self._h)
        try:
            fn()
        finally:
            N.call("ob_batch_capture_end", self._h)

    def replay(self, count=1):
        N.call("ob_batch_replay", self._h, int(count))


class _BatchBunch:
    """What the lattice nodes see: teapot_base maps queue on the batch.

    It refers to the batch weakly: the batch is then freed by reference counting, before
    the bunches and calculators it holds, never by the cycle collector in arbitrary order.
    """

    def __init__(self, batch):
        self._batch = weakref.proxy(batch)
        self._h = None  # never passed to a single-bunch entry point

    def getSyncParticle(self):
        return self._batch.bunches[0].getSyncParticle()

    def getSize(self):
        return sum(b.getSize() for b in self._batch.bunches)

    def mass(self):
        return self._batch.bunches[0].mass()

    def charge(self):
        return self._batch.bunches[0].charge()

    # the turn counter of TEAPOT_Ring (teapot.py) counts on every bunch
    def hasBunchAttrInt(self, name):
        return self._batch.bunches[0].hasBunchAttrInt(name)

    def bunchAttrInt(self, name, value=None):
        if value is not None:
            for b in self._batch.bunches:
                b.bunchAttrInt(name, value)
        return self._batch.bunches[0].bunchAttrInt(name)


class _BatchCalc:
    def __init__(self, batch):
        self._batch = weakref.proxy(batch)

    def trackBunch(self, bunch, length, pipe_radius):
        if bunch is not self._batch._proxy:
            raise ValueError("the batch calculator tracks its own batch only")
        N.call("ob_batch_sc2p5drb_track", self._batch._h, float(length), float(pipe_radius))
