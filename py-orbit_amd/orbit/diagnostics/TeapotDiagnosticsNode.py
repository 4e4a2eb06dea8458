"""Diagnostics nodes for TEAPOT lattices (reference:
py/orbit/diagnostics/TeapotDiagnosticsNode.py:23-245).  Zero-length drift nodes whose
track() writes the bunch statistics instead of moving the bunch."""
from orbit.teapot import DriftTEAPOT
from orbit.diagnostics.diagnostics import (StatLats, StatLatsSetMember, Moments,
                                           MomentsSetMember, BPMSignal)


class _DiagNode(DriftTEAPOT):
    def __init__(self, name, typename):
        DriftTEAPOT.__init__(self, name)
        self.setType(typename)
        self.setLength(0.0)
        self.position = 0.0
        self.lattlength = 0.0

    def setPosition(self, pos):
        self.position = pos

    def setLatticeLength(self, lattlength):
        self.lattlength = lattlength


class _SetMember:
    active = True

    def activate(self):
        self.active = True

    def deactivate(self):
        self.active = False


class TeapotStatLatsNode(_DiagNode):
    """Writes a StatLats line each time the bunch passes (:23-56).  The file is started
    empty, as the reference's node truncates it."""

    def __init__(self, filename, name="statlats no name"):
        _DiagNode.__init__(self, name, "statlats teapot")
        open(filename, "w").close()
        self.statlats = StatLats(filename)

    def track(self, paramsDict):
        self.statlats.writeStatLats(self.position, paramsDict["bunch"], self.lattlength)

    def closeStatLats(self):
        self.statlats.closeStatLats()


class TeapotStatLatsNodeSetMember(_SetMember, _DiagNode):
    """A StatLats node of a set sharing one file (:58-99)."""

    def __init__(self, file, name="statlats no name"):
        _DiagNode.__init__(self, name, "statlats teapot")
        self.statlats = StatLatsSetMember(file)
        self.active = True
        self.file = file

    def track(self, paramsDict):
        if self.active:
            self.statlats.writeStatLats(self.position, paramsDict["bunch"], self.lattlength)

    def resetFile(self, file):
        self.file = file
        self.statlats.resetFile(file)


class TeapotMomentsNode(_DiagNode):
    """Writes a Moments line each time the bunch passes (:102-133)."""

    def __init__(self, filename, order, nodispersion=True, emitnorm=False,
                 name="moments no name"):
        _DiagNode.__init__(self, name, "moments teapot")
        open(filename, "w").close()
        self.moments = Moments(filename, order, nodispersion, emitnorm)

    def track(self, paramsDict):
        self.moments.writeMoments(self.position, paramsDict["bunch"], self.lattlength)

    def closeMoments(self):
        self.moments.closeMoments()


class TeapotMomentsNodeSetMember(_SetMember, _DiagNode):
    """A Moments node of a set sharing one file (:136-178)."""

    def __init__(self, file, order, nodispersion=True, emitnorm=False,
                 name="moments no name"):
        _DiagNode.__init__(self, str(name), "moments teapot")
        self.file = file
        self.moments = MomentsSetMember(file, order, nodispersion, emitnorm)
        self.active = True

    def track(self, paramsDict):
        if self.active:
            self.moments.writeMoments(self.position, paramsDict["bunch"], self.lattlength)

    def resetFile(self, file):
        self.file = file
        self.moments.resetFile(file)


class TeapotBPMSignalNode(_DiagNode):
    """Records the bunch's average x, y as a beam position monitor (:212-245)."""

    def __init__(self, name="BPMSignal no name"):
        _DiagNode.__init__(self, name, "BPMSignal")
        self.bpm = BPMSignal()

    def track(self, paramsDict):
        self.bpm.analyzeSignal(paramsDict["bunch"])

    def getSignal(self):
        return self.bpm.getSignalX(), self.bpm.getSignalY()
