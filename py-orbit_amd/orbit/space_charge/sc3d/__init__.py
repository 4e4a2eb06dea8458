"""3D space-charge nodes (reference: py/orbit/space_charge/sc3d)."""
from orbit.space_charge.sc3d.scAccNodes import SC3D_AccNode
from orbit.space_charge.sc3d.scLatticeModifications import setSC3DAccNodes

__all__ = ["SC3D_AccNode", "setSC3DAccNodes"]
