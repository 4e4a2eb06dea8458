"""setSC3DAccNodes (reference: py/orbit/space_charge/sc3d/scLatticeModifications.py:14-26)."""
from orbit.space_charge.sc3d.scAccNodes import SC3D_AccNode
from orbit.space_charge.scLatticeModifications import setSC_General_AccNodes


def setSC3DAccNodes(lattice, sc_path_length_min, space_charge_calculator):
    scNodes = setSC_General_AccNodes(lattice, sc_path_length_min, space_charge_calculator,
                                     SC3D_AccNode)
    for scNode in scNodes:
        scNode.setName(scNode.getName() + "SC3D")
    lattice.initialize()
    return scNodes
