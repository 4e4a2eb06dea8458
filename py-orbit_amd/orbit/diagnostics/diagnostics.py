"""Statistical writers (reference: py/orbit/diagnostics/diagnostics.py:20-248).

One process holds the whole bunch, so every writer writes (the reference writes on MPI
rank 0 only).  Lines keep the reference's columns and tab separators; numbers are
written with Python's str()."""
from bunch import BunchTwissAnalysis
from orbit.utils.consts import speed_of_light


def _time(bunch, lattlength):
    """Synchronous-particle time, in turns when the lattice length is given
    (diagnostics.py:44-47)."""
    sp = bunch.getSyncParticle()
    t = sp.time()
    if lattlength > 0:
        t = sp.time() / (lattlength / (sp.beta() * speed_of_light))
    return t


def _statlats_line(twiss, s, bunch, lattlength):
    twiss.analyzeBunch(bunch)
    cols = (s, _time(bunch, lattlength), twiss.getEmittance(0), twiss.getEmittance(1),
            twiss.getBeta(0), twiss.getBeta(1), twiss.getAlpha(0), twiss.getAlpha(1),
            twiss.getDispersion(0), twiss.getDispersionDerivative(0))
    return "\t".join(str(c) for c in cols) + "\n"


def _moments_line(twiss, order, dispterm, emitnormterm, s, bunch, lattlength):
    t = _time(bunch, lattlength)
    twiss.computeBunchMoments(bunch, order, dispterm, emitnormterm)
    out = [str(s), str(t)]
    for i in range(order + 1):
        for j in range(i + 1):
            out.append(str(twiss.getBunchMoment(i - j, j)))
    return "\t".join(out) + "\t\n"


class StatLats:
    """Appends s, time, emittances, betas, alphas, x dispersion and its derivative to a
    file per call (diagnostics.py:20-61)."""

    def __init__(self, filename):
        self.file_out = open(filename, "a")
        self.bunchtwissanalysis = BunchTwissAnalysis()

    def writeStatLats(self, s, bunch, lattlength=0):
        self.file_out.write(_statlats_line(self.bunchtwissanalysis, s, bunch, lattlength))

    def closeStatLats(self):
        self.file_out.close()


class StatLatsSetMember:
    """StatLats writing into a file shared by a node set (diagnostics.py:64-108)."""

    def __init__(self, file):
        self.file_out = file
        self.bunchtwissanalysis = BunchTwissAnalysis()

    def writeStatLats(self, s, bunch, lattlength=0):
        self.file_out.write(_statlats_line(self.bunchtwissanalysis, s, bunch, lattlength))

    def closeStatLats(self):
        self.file_out.close()

    def resetFile(self, file):
        self.file_out = file


def _flags(nodispersion, emitnorm):
    # diagnostics.py:118-126: nodispersion=True subtracts the dispersive part of x
    return (1 if nodispersion else -1), (1 if emitnorm else -1)


class Moments:
    """Appends s, time and the x-y moments <x^(i-j) y^j>, i = 0..order, j = 0..i, per call
    (diagnostics.py:112-151)."""

    def __init__(self, filename, order, nodispersion, emitnorm):
        self.file_out = open(filename, "a")
        self.bunchtwissanalysis = BunchTwissAnalysis()
        self.order = order
        self.dispterm, self.emitnormterm = _flags(nodispersion, emitnorm)

    def writeMoments(self, s, bunch, lattlength=0):
        self.file_out.write(_moments_line(self.bunchtwissanalysis, self.order, self.dispterm,
                                          self.emitnormterm, s, bunch, lattlength))

    def closeMoments(self):
        self.file_out.close()


class MomentsSetMember:
    """Moments writing into a file shared by a node set (diagnostics.py:155-199)."""

    def __init__(self, file, order, nodispersion, emitnorm):
        self.file_out = file
        self.order = order
        self.bunchtwissanalysis = BunchTwissAnalysis()
        self.dispterm, self.emitnormterm = _flags(nodispersion, emitnorm)

    def writeMoments(self, s, bunch, lattlength=0):
        self.file_out.write(_moments_line(self.bunchtwissanalysis, self.order, self.dispterm,
                                          self.emitnormterm, s, bunch, lattlength))

    def resetFile(self, file):
        self.file_out = file


class BPMSignal:
    """Average x, xp, y, yp of the bunch (diagnostics.py:201-248)."""

    def __init__(self):
        self.bunchtwissanalysis = BunchTwissAnalysis()
        self.xAvg = self.yAvg = self.xpAvg = self.ypAvg = 0.0

    def analyzeSignal(self, bunch):
        t = self.bunchtwissanalysis
        t.analyzeBunch(bunch)
        self.xAvg, self.xpAvg = t.getAverage(0), t.getAverage(1)
        self.yAvg, self.ypAvg = t.getAverage(2), t.getAverage(3)

    def getSignalX(self):
        return self.xAvg

    def getSignalXP(self):
        return self.xpAvg

    def getSignalY(self):
        return self.yAvg

    def getSignalYP(self):
        return self.ypAvg
