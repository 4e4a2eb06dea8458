"""Space-charge lattice nodes (reference: py/orbit/space_charge)."""
from orbit.space_charge.scAccNodes import SC_Base_AccNode
from orbit.space_charge.scLatticeModifications import setSC_General_AccNodes

__all__ = ["SC_Base_AccNode", "setSC_General_AccNodes"]
