"""Diagnostics nodes and profiles (reference: py/orbit/diagnostics).

The statistics they write come from the device: BunchTwissAnalysis reads the bunch's
moment sums, computeBunchMoments its x-y power sums, profiles bins the bunch on the
device, and BunchExtremaCalculator finds the extrema there; no node downloads the bunch."""
from orbit.diagnostics.diagnostics import (StatLats, StatLatsSetMember, Moments,
                                           MomentsSetMember, BPMSignal)
from orbit.diagnostics.profiles import profiles
from orbit.diagnostics.diagnosticsLatticeModifications import (
    addTeapotDiagnosticsNode, addTeapotDiagnosticsNodeAsChild, addTeapotStatLatsNodeSet,
    addTeapotMomentsNodeSet, diagnosticsNodeSetController)
from orbit.diagnostics.TeapotDiagnosticsNode import (TeapotStatLatsNode,
                                                     TeapotStatLatsNodeSetMember,
                                                     TeapotMomentsNode,
                                                     TeapotMomentsNodeSetMember,
                                                     TeapotBPMSignalNode)

__all__ = ["StatLats", "StatLatsSetMember", "TeapotStatLatsNode",
           "TeapotStatLatsNodeSetMember", "Moments", "MomentsSetMember", "TeapotMomentsNode",
           "TeapotMomentsNodeSetMember", "addTeapotDiagnosticsNode",
           "addTeapotDiagnosticsNodeAsChild", "addTeapotStatLatsNodeSet",
           "addTeapotMomentsNodeSet", "diagnosticsNodeSetController", "TeapotBPMSignalNode",
           "BPMSignal", "profiles"]
