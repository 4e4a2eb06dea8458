"""Base space-charge node (reference: py/orbit/space_charge/scAccNodes.py)."""
from orbit.lattice import AccNodeBunchTracker


class SC_Base_AccNode(AccNodeBunchTracker):
    def __init__(self, sc_calculator, name="no name"):
        AccNodeBunchTracker.__init__(self, name)
        self.setType("SC_Base")
        self.sc_length = 0.0
        self.switcher = True
        self.sc_calculator = sc_calculator

    def isRFGap(self):
        return False

    def setLengthOfSC(self, sc_length):
        """Path length used by the kick (the distance to the next SC node)."""
        self.sc_length = sc_length

    def getLengthOfSC(self):
        return self.sc_length

    def setCalculationOn(self, switcher):
        # the reference sets True whatever the argument (scAccNodes.py:47-51)
        self.switcher = True

    def getCalculationOn(self):
        return self.switcher

    def getCalculator(self):
        return self.sc_calculator

    def trackDesign(self, paramsDict):
        pass

    def track(self, paramsDict):
        pass
