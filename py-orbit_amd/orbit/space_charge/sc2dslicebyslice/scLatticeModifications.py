"""setSC2DSliceBySliceAccNodes (reference: py/orbit/space_charge/sc2dslicebyslice/
scLatticeModifications.py:17-31)."""
from orbit.space_charge.sc2dslicebyslice.scAccNodes import SC2DSliceBySlice_AccNode
from orbit.space_charge.scLatticeModifications import setSC_General_AccNodes


def setSC2DSliceBySliceAccNodes(lattice, sc_path_length_min, space_charge_calculator,
                                boundary=None):
    scNodes = setSC_General_AccNodes(lattice, sc_path_length_min, space_charge_calculator,
                                     SC2DSliceBySlice_AccNode)
    for scNode in scNodes:
        scNode.setName(scNode.getName() + "SC2DSliceBySlice")
        scNode.setBoundary(boundary)
    lattice.initialize()
    return scNodes
