"""TEAPOT lattice elements (reference: py/orbit/teapot/teapot.py).

Each element's track() issues the same teapot_base map calls, in the same order and with
the same arguments, as the reference element; the maps run on the device.  Apertures
(which need the reference's aperture module) are not built.
"""
import os

from orbit.lattice import AccActionsContainer, AccLattice, AccNode, AccNodeBunchTracker
from orbit.teapot_base import TPB
from orbit.utils import orbitFinalize


class TEAPOT_Lattice(AccLattice):
    """Lattice of TEAPOT nodes (teapot.py:43-151)."""

    def __init__(self, name="no name"):
        AccLattice.__init__(self, name)
        self.useCharge = 1

    def _addChildren(self):
        pass

    def trackBunch(self, bunch, paramsDict=None, actionContainer=None, index_start=-1,
                   index_stop=-1):
        """Track the bunch through nodes index_start..index_stop inclusive (:127-141)."""
        if paramsDict is None:
            paramsDict = {}
        if actionContainer is None:
            actionContainer = AccActionsContainer("Bunch Tracking")
        paramsDict["bunch"] = bunch
        paramsDict["useCharge"] = self.useCharge

        def track(params):
            params["node"].track(params)

        actionContainer.addAction(track, AccActionsContainer.BODY)
        self.trackActions(actionContainer, paramsDict, index_start, index_stop)
        actionContainer.removeAction(track, AccActionsContainer.BODY)

    def setUseRealCharge(self, useCharge=1):
        self.useCharge = useCharge

    def getUseRealCharge(self):
        return self.useCharge


class TEAPOT_Ring(TEAPOT_Lattice):
    """Ring: every first-level node with length gets a bunch-wrap EXIT child (:154-175)."""

    def __init__(self, name="no name"):
        TEAPOT_Lattice.__init__(self, name)

    def _addChildren(self):
        TEAPOT_Lattice.initialize(self)
        for node in self.getNodes():
            if node.getLength() > 0.0:
                wrap = BunchWrapTEAPOT(node.getName() + ":Bunch_Wrap:Exit")
                wrap.getParamsDict()["ring_length"] = self.getLength()
                node.addChildNode(wrap, AccNode.EXIT)
        self.getNodes().append(TurnCounterTEAPOT())


class BaseTEAPOT(AccNodeBunchTracker):
    def __init__(self, name="no name"):
        AccNodeBunchTracker.__init__(self, name)
        self.setType("base teapot")


class TurnCounterTEAPOT(BaseTEAPOT):
    """Increments the bunch's integer attribute TurnNumber if present (:413-428)."""

    def __init__(self, name="TurnCounter"):
        BaseTEAPOT.__init__(self, name)
        self.setType("turn counter")

    def track(self, paramsDict):
        bunch = paramsDict["bunch"]
        if bunch.hasBunchAttrInt("TurnNumber") != 0:
            bunch.bunchAttrInt("TurnNumber", bunch.bunchAttrInt("TurnNumber") + 1)


class TiltTEAPOT(BaseTEAPOT):
    """Rotation about the axis at an element's ends (:1298-1329)."""

    def __init__(self, name="tilt no name", angle=0.0):
        AccNode.__init__(self, name)
        self.__angle = angle
        self.setType("tilt teapot")

    def setTiltAngle(self, angle=0.0):
        self.__angle = angle

    def getTiltAngle(self):
        return self.__angle

    def track(self, paramsDict):
        if self.__angle != 0.0:
            TPB.rotatexy(paramsDict["bunch"], self.__angle)


class FringeFieldTEAPOT(BaseTEAPOT):
    """Entrance/exit fringe node sharing its parent's parameters (:1331-1377)."""

    def __init__(self, parentNode, trackFunction=None, name="fringe field no name"):
        AccNode.__init__(self, name)
        self.setParamsDict(parentNode.getParamsDict())
        self.__trackFunc = trackFunction
        self.__usage = True
        self.setType("fringeField teapot")

    def track(self, paramsDict):
        if self.__trackFunc is not None:
            self.__trackFunc(self, paramsDict)

    def setFringeFieldFunction(self, trackFunction=None):
        self.__trackFunc = trackFunction

    def getFringeFieldFunction(self):
        return self.__trackFunc

    def setUsage(self, usage=True):
        self.__usage = usage

    def getUsage(self):
        return self.__usage


class NodeTEAPOT(BaseTEAPOT):
    """Real TEAPOT element: tilt in, fringe in / body / fringe out, tilt out (:430-546)."""

    def __init__(self, name="no name"):
        BaseTEAPOT.__init__(self, name)
        self.__tiltIN = TiltTEAPOT()
        self.__tiltOUT = TiltTEAPOT()
        self.__fringeIN = FringeFieldTEAPOT(self)
        self.__fringeOUT = FringeFieldTEAPOT(self)
        self.__tiltIN.setName(name + "_tilt_in")
        self.__tiltOUT.setName(name + "_tilt_out")
        self.__fringeIN.setName(name + "_fringe_in")
        self.__fringeOUT.setName(name + "_fringe_out")
        self.addChildNode(self.__tiltIN, AccNode.ENTRANCE)
        self.addChildNode(self.__fringeIN, AccNode.ENTRANCE)
        self.addChildNode(self.__fringeOUT, AccNode.EXIT)
        self.addChildNode(self.__tiltOUT, AccNode.EXIT)
        self.addParam("tilt", self.__tiltIN.getTiltAngle())
        self.setType("node teapot")

    def setTiltAngle(self, angle=0.0):
        self.setParam("tilt", angle)
        self.__tiltIN.setTiltAngle(angle)
        self.__tiltOUT.setTiltAngle(-1.0 * angle)

    def getTiltAngle(self):
        return self.__tiltIN.getTiltAngle()

    def getNodeFringeFieldIN(self):
        return self.__fringeIN

    def getNodeFringeFieldOUT(self):
        return self.__fringeOUT

    def getNodeTiltIN(self):
        return self.__tiltIN

    def getNodeTiltOUT(self):
        return self.__tiltOUT

    def setFringeFieldFunctionIN(self, trackFunction=None):
        self.__fringeIN.setFringeFieldFunction(trackFunction)

    def setFringeFieldFunctionOUT(self, trackFunction=None):
        self.__fringeOUT.setFringeFieldFunction(trackFunction)

    def getFringeFieldFunctionIN(self):
        return self.__fringeIN.getFringeFieldFunction()

    def getFringeFieldFunctionOUT(self):
        return self.__fringeOUT.getFringeFieldFunction()

    def setUsageFringeFieldIN(self, usage=True):
        self.__fringeIN.setUsage(usage)

    def getUsageFringeFieldIN(self):
        return self.__fringeIN.getUsage()

    def setUsageFringeFieldOUT(self, usage=True):
        self.__fringeOUT.setUsage(usage)

    def getUsageFringeFieldOUT(self):
        return self.__fringeOUT.getUsage()


def _use_charge(paramsDict):
    return paramsDict.get("useCharge", 1)


def _split_parts(node, what):
    """Parts 0 and n-1 get half a step, inner parts a full step (:754-773, :904-923)."""
    nParts = node.getnParts()
    if nParts < 2:
        orbitFinalize("The " + what + " TEAPOT class instance should have more than 2 parts!"
                      + os.linesep + "Name of element=" + node.getName()
                      + os.linesep + "nParts =" + str(nParts))
    half = (node.getLength() / (nParts - 1)) / 2.0
    node.setLength(half, 0)
    node.setLength(half, nParts - 1)
    for i in range(nParts - 2):
        node.setLength(half + half, i + 1)


class DriftTEAPOT(NodeTEAPOT):
    """Drift (:548-565)."""

    def __init__(self, name="drift no name"):
        NodeTEAPOT.__init__(self, name)
        self.setType("drift teapot")

    def track(self, paramsDict):
        TPB.drift(paramsDict["bunch"], self.getLength(self.getActivePartIndex()))


class BunchWrapTEAPOT(NodeTEAPOT):
    """Longitudinal wrap into +-ring_length/2 (:626-645)."""

    def __init__(self, name="drift no name"):
        NodeTEAPOT.__init__(self, name)
        self.setType("bunch_wrap_teapot")
        self.addParam("ring_length", 0.0)

    def track(self, paramsDict):
        TPB.wrapbunch(paramsDict["bunch"], self.getParam("ring_length"))


class MultipoleTEAPOT(NodeTEAPOT):
    """Multipole: drift halves around thin kicks, hard-edge fringes (:685-816)."""

    def __init__(self, name="multipole no name"):
        NodeTEAPOT.__init__(self, name)
        self.addParam("poles", [])
        self.addParam("kls", [])
        self.addParam("skews", [])
        self.setnParts(2)
        self.waveform = None
        owner = self

        def fringe(out):
            def f(node, paramsDict):
                if not node.getUsage():
                    return
                length = paramsDict["parentNode"].getLength()
                if length == 0.0:
                    return
                strength = owner.waveform.getStrength() if owner.waveform else 1.0
                bunch = paramsDict["bunch"]
                uc = _use_charge(paramsDict)
                fn = TPB.multpfringeOUT if out else TPB.multpfringeIN
                for pole, kl, skew in zip(node.getParam("poles"), node.getParam("kls"),
                                          node.getParam("skews")):
                    fn(bunch, pole, strength * kl / length, skew, uc)
            return f

        self.setFringeFieldFunctionIN(fringe(False))
        self.setFringeFieldFunctionOUT(fringe(True))
        self.setType("multipole teapot")

    def initialize(self):
        _split_parts(self, "Multipole")

    def track(self, paramsDict):
        nParts = self.getnParts()
        index = self.getActivePartIndex()
        length = self.getLength(index)
        strength = self.waveform.getStrength() if self.waveform else 1.0
        bunch = paramsDict["bunch"]
        uc = _use_charge(paramsDict)
        if index == 0:
            TPB.drift(bunch, length)
            return
        for pole, kl, skew in zip(self.getParam("poles"), self.getParam("kls"),
                                  self.getParam("skews")):
            TPB.multp(bunch, pole, strength * kl / (nParts - 1), skew, uc)
        TPB.drift(bunch, length)

    def setWaveform(self, waveform):
        self.waveform = waveform


class QuadTEAPOT(NodeTEAPOT):
    """Quadrupole with optional multipoles and hard-edge fringes (:819-976)."""

    def __init__(self, name="quad no name"):
        NodeTEAPOT.__init__(self, name)
        self.addParam("kq", 0.0)
        self.addParam("poles", [])
        self.addParam("kls", [])
        self.addParam("skews", [])
        self.setnParts(2)
        self.waveform = None
        owner = self

        def fringe(out):
            def f(node, paramsDict):
                if not node.getUsage():
                    return
                strength = owner.waveform.getStrength() if owner.waveform else 1.0
                kq = strength * node.getParam("kq")
                length = paramsDict["parentNode"].getLength()
                bunch = paramsDict["bunch"]
                uc = _use_charge(paramsDict)
                (TPB.quadfringeOUT if out else TPB.quadfringeIN)(bunch, kq, uc)
                if length == 0.0:
                    return
                fn = TPB.multpfringeOUT if out else TPB.multpfringeIN
                for pole, kl, skew in zip(node.getParam("poles"), node.getParam("kls"),
                                          node.getParam("skews")):
                    fn(bunch, pole, strength * kl / length, skew, uc)
            return f

        self.setFringeFieldFunctionIN(fringe(False))
        self.setFringeFieldFunctionOUT(fringe(True))
        self.getNodeTiltIN().setType("quad tilt in")
        self.getNodeTiltOUT().setType("quad tilt out")
        self.getNodeFringeFieldIN().setType("quad fringe in")
        self.getNodeFringeFieldOUT().setType("quad fringe out")
        self.setType("quad teapot")

    def initialize(self):
        _split_parts(self, "Quad Combined Function")

    def track(self, paramsDict):
        nParts = self.getnParts()
        index = self.getActivePartIndex()
        length = self.getLength(index)
        strength = self.waveform.getStrength() if self.waveform else 1.0
        kq = strength * self.getParam("kq")
        bunch = paramsDict["bunch"]
        uc = _use_charge(paramsDict)
        if index == 0:
            TPB.quad1(bunch, length, kq, uc)
            return
        half = length / 2.0 if index < nParts - 1 else length
        TPB.quad2(bunch, half)
        for pole, kl, skew in zip(self.getParam("poles"), self.getParam("kls"),
                                  self.getParam("skews")):
            TPB.multp(bunch, pole, strength * kl / (nParts - 1), skew, uc)
        TPB.quad2(bunch, half)
        TPB.quad1(bunch, length, kq, uc)

    def setWaveform(self, waveform):
        self.waveform = waveform


class KickTEAPOT(NodeTEAPOT):
    """Distributed kicker: drift + kick per part, last part drift only (:1221-1296)."""

    def __init__(self, name="kick no name"):
        NodeTEAPOT.__init__(self, name)
        self.addParam("kx", 0.0)
        self.addParam("ky", 0.0)
        self.addParam("dE", 0.0)
        self.setType("kick teapot")
        self.setnParts(2)
        self.waveform = None

    def initialize(self):
        _split_parts(self, "Kick")

    def track(self, paramsDict):
        nParts = self.getnParts()
        index = self.getActivePartIndex()
        length = self.getLength(index)
        strength = self.waveform.getStrength() if self.waveform else 1.0
        kx = strength * self.getParam("kx") / (nParts - 1)
        ky = strength * self.getParam("ky") / (nParts - 1)
        dE = self.getParam("dE") / (nParts - 1)
        bunch = paramsDict["bunch"]
        TPB.drift(bunch, length)
        if index < nParts - 1:
            TPB.kick(bunch, kx, ky, dE, _use_charge(paramsDict))

    def setWaveform(self, waveform):
        self.waveform = waveform


class BendTEAPOT(NodeTEAPOT):
    """Combined-function bend: wedge edges and bend fringes around the bend1..4 splitting
    (:978-1150)."""

    def __init__(self, name="bend no name"):
        NodeTEAPOT.__init__(self, name)
        self.addParam("poles", [])
        self.addParam("kls", [])
        self.addParam("skews", [])
        self.addParam("ea1", 0.0)
        self.addParam("ea2", 0.0)
        self.addParam("rho", 0.0)
        self.addParam("theta", 1.0e-36)
        self.setnParts(2)

        def multipole_fringes(node, paramsDict, fn, length):
            if length == 0.0:
                return
            bunch = paramsDict["bunch"]
            uc = _use_charge(paramsDict)
            for pole, kl, skew in zip(node.getParam("poles")[:], node.getParam("kls")[:],
                                      node.getParam("skews")[:]):
                fn(bunch, pole, kl / length, skew, uc)

        def fringeIN(node, paramsDict):
            e = node.getParam("ea1")
            rho = node.getParam("rho")
            parent = paramsDict["parentNode"]
            length = parent.getLength()
            bunch = paramsDict["bunch"]
            if e != 0.0:
                TPB.wedgedrift(bunch, e, 0)
                if node.getUsage():
                    TPB.wedgerotate(bunch, e, 0)
                    TPB.bendfringeIN(bunch, rho)
                    multipole_fringes(node, paramsDict, TPB.multpfringeIN, length)
                    TPB.wedgerotate(bunch, e, 1)
                poles = node.getParam("poles")[:]
                TPB.wedgebendCF(bunch, e, 0, rho, len(poles), poles, node.getParam("kls")[:],
                                node.getParam("skews")[:], parent.getnParts() - 1,
                                _use_charge(paramsDict))
            elif node.getUsage():
                TPB.bendfringeIN(bunch, rho)
                multipole_fringes(node, paramsDict, TPB.multpfringeIN, length)

        def fringeOUT(node, paramsDict):
            e = node.getParam("ea2")
            rho = node.getParam("rho")
            parent = paramsDict["parentNode"]
            length = parent.getLength()
            bunch = paramsDict["bunch"]
            if e != 0.0:
                poles = node.getParam("poles")[:]
                TPB.wedgebendCF(bunch, e, 1, rho, len(poles), poles, node.getParam("kls")[:],
                                node.getParam("skews")[:], parent.getnParts() - 1,
                                _use_charge(paramsDict))
                if node.getUsage():
                    TPB.wedgerotate(bunch, -e, 0)
                    TPB.bendfringeOUT(bunch, rho)
                    multipole_fringes(node, paramsDict, TPB.multpfringeOUT, length)
                    TPB.wedgerotate(bunch, -e, 1)
                TPB.wedgedrift(bunch, e, 1)
            elif node.getUsage():
                TPB.bendfringeOUT(bunch, rho)
                multipole_fringes(node, paramsDict, TPB.multpfringeOUT, length)

        self.setFringeFieldFunctionIN(fringeIN)
        self.setFringeFieldFunctionOUT(fringeOUT)
        self.setType("bend teapot")

    def initialize(self):
        if self.getnParts() < 2:
            orbitFinalize("The Bend Combined Functions TEAPOT class instance should have more "
                          "than 2 parts!" + os.linesep + "Method initialize():" + os.linesep
                          + "Name of element=" + self.getName() + os.linesep
                          + "Type of element=" + self.getType() + os.linesep
                          + "nParts =" + str(self.getnParts()))
        self.addParam("rho", self.getLength() / self.getParam("theta"))
        _split_parts(self, "Bend Combined Functions")

    def track(self, paramsDict):
        nParts = self.getnParts()
        index = self.getActivePartIndex()
        length = self.getLength(index)
        bunch = paramsDict["bunch"]
        uc = _use_charge(paramsDict)
        theta = self.getParam("theta") / (nParts - 1)
        if index == 0:
            TPB.bend1(bunch, length, theta / 2.0)
            return
        last = index == nParts - 1
        step = length if last else length / 2.0
        TPB.bend2(bunch, step)
        TPB.bend3(bunch, theta / 2.0)
        TPB.bend4(bunch, theta / 2.0)
        for pole, kl, skew in zip(self.getParam("poles"), self.getParam("kls"),
                                  self.getParam("skews")):
            TPB.multp(bunch, pole, kl / (nParts - 1), skew, uc)
        TPB.bend4(bunch, theta / 2.0)
        TPB.bend3(bunch, theta / 2.0)
        TPB.bend2(bunch, step)
        TPB.bend1(bunch, length, theta / 2.0 if last else theta)


class SolenoidTEAPOT(NodeTEAPOT):
    """Solenoid (:648-682); B in 1/m, scaled by an optional waveform."""

    def __init__(self, name="solenoid no name"):
        NodeTEAPOT.__init__(self, name)
        self.setType("solenoid teapot")
        self.addParam("B", 0.0)
        self.waveform = None

    def track(self, paramsDict):
        length = self.getLength(self.getActivePartIndex())
        strength = self.waveform.getStrength() if self.waveform else 1.0
        TPB.soln(paramsDict["bunch"], length, strength * self.getParam("B"),
                 _use_charge(paramsDict))

    def setWaveform(self, waveform):
        self.waveform = waveform


class RingRFTEAPOT(NodeTEAPOT):
    """Ring RF cavity: one RingRF call per harmonic (:1153-1219)."""

    def __init__(self, name="RingRF no name"):
        NodeTEAPOT.__init__(self, name)
        self.addParam("harmonics", [])
        self.addParam("voltages", [])
        self.addParam("phases", [])
        self.addParam("ring_length", 0.0)
        self.setType("RingRF teapot")
        self.setnParts(1)

    def addRF(self, harmonic, voltage, phase):
        self.getParam("harmonics").append(harmonic)
        self.getParam("voltages").append(voltage)
        self.getParam("phases").append(phase)

    def initialize(self):
        if self.getnParts() > 1:
            orbitFinalize("The Ring RF TEAPOT class instance should not have more than 1 part!"
                          + os.linesep + "Method initialize():" + os.linesep
                          + "Name of element=" + self.getName() + os.linesep
                          + "Type of element=" + self.getType() + os.linesep
                          + "nParts =" + str(self.getnParts()) + os.linesep
                          + "length =" + str(self.getLength()))

    def track(self, paramsDict):
        bunch = paramsDict["bunch"]
        uc = _use_charge(paramsDict)
        ring_length = self.getParam("ring_length")
        for h, v, ph in zip(self.getParam("harmonics"), self.getParam("voltages"),
                            self.getParam("phases")):
            TPB.RingRF(bunch, ring_length, h, v, ph, uc)


class MonitorTEAPOT(NodeTEAPOT):
    """Beam-position monitor: records the bunch centroid (:601-623), from the device
    moments of BunchTwissAnalysis."""

    def __init__(self, name="Monitor no name"):
        from bunch import BunchTwissAnalysis
        NodeTEAPOT.__init__(self, name)
        self.setType("monitor teapot")
        self.twiss = BunchTwissAnalysis()

    def track(self, paramsDict):
        self.twiss.analyzeBunch(paramsDict["bunch"])
        self.addParam("xAvg", self.twiss.getAverage(0))
        self.addParam("xpAvg", self.twiss.getAverage(1))
        self.addParam("yAvg", self.twiss.getAverage(2))
        self.addParam("ypAvg", self.twiss.getAverage(3))
