"""TEAPOT elements (reference: py/orbit/teapot/__init__.py)."""
from orbit.teapot.teapot import (TEAPOT_Lattice, TEAPOT_Ring, BaseTEAPOT, NodeTEAPOT, DriftTEAPOT,
                                 BunchWrapTEAPOT, QuadTEAPOT, MultipoleTEAPOT, KickTEAPOT,
                                 TiltTEAPOT, FringeFieldTEAPOT, TurnCounterTEAPOT, BendTEAPOT,
                                 SolenoidTEAPOT, RingRFTEAPOT, MonitorTEAPOT, TPB)

__all__ = ["TEAPOT_Lattice", "TEAPOT_Ring", "BaseTEAPOT", "NodeTEAPOT", "DriftTEAPOT",
           "BunchWrapTEAPOT", "QuadTEAPOT", "MultipoleTEAPOT", "KickTEAPOT", "TiltTEAPOT",
           "FringeFieldTEAPOT", "TurnCounterTEAPOT", "BendTEAPOT", "SolenoidTEAPOT",
           "RingRFTEAPOT", "MonitorTEAPOT", "TPB"]
