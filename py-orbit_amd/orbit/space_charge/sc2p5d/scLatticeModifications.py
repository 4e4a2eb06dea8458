"""setSC2p5DAccNodes and setSC2p5DrbAccNodes (reference: py/orbit/space_charge/sc2p5d/
scLatticeModifications.py:17-47)."""
from orbit.space_charge.sc2p5d.scAccNodes import SC2p5D_AccNode, SC2p5Drb_AccNode
from orbit.space_charge.scLatticeModifications import setSC_General_AccNodes


def setSC2p5DAccNodes(lattice, sc_path_length_min, space_charge_calculator, boundary=None):
    scNodes = setSC_General_AccNodes(lattice, sc_path_length_min, space_charge_calculator,
                                     SC2p5D_AccNode)
    for scNode in scNodes:
        scNode.setName(scNode.getName() + "SC2p5D")
        scNode.setBoundary(boundary)
    lattice.initialize()
    return scNodes


def setSC2p5DrbAccNodes(lattice, sc_path_length_min, space_charge_calculator, pipe_radius):
    scNodes = setSC_General_AccNodes(lattice, sc_path_length_min, space_charge_calculator,
                                     SC2p5Drb_AccNode)
    for scNode in scNodes:
        scNode.setName(scNode.getName() + "SC2p5Drb")
        scNode.setPipeRadius(pipe_radius)
    lattice.initialize()
    return scNodes
