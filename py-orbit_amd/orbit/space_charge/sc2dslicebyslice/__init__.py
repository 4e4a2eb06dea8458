"""2D slice-by-slice space-charge nodes (reference: py/orbit/space_charge/sc2dslicebyslice)."""
from orbit.space_charge.sc2dslicebyslice.scAccNodes import SC2DSliceBySlice_AccNode
from orbit.space_charge.sc2dslicebyslice.scLatticeModifications import (
    setSC2DSliceBySliceAccNodes)

__all__ = ["SC2DSliceBySlice_AccNode", "setSC2DSliceBySliceAccNodes"]
