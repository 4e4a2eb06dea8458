"""SC2DSliceBySlice_AccNode (reference: py/orbit/space_charge/sc2dslicebyslice/
scAccNodes.py:18-51)."""
from orbit.space_charge.scAccNodes import SC_Base_AccNode


class SC2DSliceBySlice_AccNode(SC_Base_AccNode):
    """Calls SpaceChargeCalcSliceBySlice2D.trackBunch(bunch, sc_length[, boundary])."""

    def __init__(self, sc_calculator, name="no name"):
        SC_Base_AccNode.__init__(self, sc_calculator, name)
        self.setType("SC2DSliceBySlice")
        self.boundary = None

    def setBoundary(self, boundary):
        self.boundary = boundary

    def getBoundary(self):
        return self.boundary

    def track(self, paramsDict):
        if self.switcher is not True:
            return
        bunch = paramsDict["bunch"]
        if self.boundary is not None:
            self.sc_calculator.trackBunch(bunch, self.sc_length, self.boundary)
        else:
            self.sc_calculator.trackBunch(bunch, self.sc_length)
