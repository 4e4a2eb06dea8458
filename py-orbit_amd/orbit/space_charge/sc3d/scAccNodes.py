"""SC3D_AccNode (reference: py/orbit/space_charge/sc3d/scAccNodes.py:18-35)."""
from orbit.space_charge.scAccNodes import SC_Base_AccNode


class SC3D_AccNode(SC_Base_AccNode):
    def __init__(self, sc_calculator, name="no name"):
        SC_Base_AccNode.__init__(self, sc_calculator, name)
        self.setType("SC3D")

    def track(self, paramsDict):
        if self.switcher is not True:
            return
        self.sc_calculator.trackBunch(paramsDict["bunch"], self.sc_length)
