"""Insertion of SC nodes into a lattice (reference:
py/orbit/space_charge/scLatticeModifications.py:10-61).

Walking the first-level nodes part by part, an SC node is placed BEFORE a part whenever
the path accumulated since the previous SC node exceeds sc_path_length_min; a first SC
node sits before part 0 of the first node.  Each SC node's length is the path to the
next SC node (the last one: the rest of the lattice).
"""
import sys

from orbit.lattice import AccNode


def setSC_General_AccNodes(lattice, sc_path_length_min, space_charge_calculator,
                           SC_NodeConstructor):
    accNodes = lattice.getNodes()
    if len(accNodes) == 0:
        return
    places = []  # (node, part index, position, path since previous SC node)
    total = 0.0
    running = 0.0
    for accNode in accNodes:
        for ip in range(accNode.getnParts()):
            part_length = accNode.getLength(ip)
            if part_length > 1.0:
                print("Warning! Node ", accNode.getName(), " has length ", part_length,
                      "m which is great than 1 m.  Space charge algorithm may be innacurate!",
                      file=sys.stderr)
            if running > sc_path_length_min:
                places.append((accNode, ip, total, running))
                running = 0.0
            running += part_length
            total += part_length
    rest_length = total - places[-1][2] if places else total
    places.insert(0, (accNodes[0], 0, 0.0, rest_length))
    scNodes = []
    for k in range(len(places) - 1):
        accNode, part_index = places[k][0], places[k][1]
        scNode = SC_NodeConstructor(space_charge_calculator,
                                    accNode.getName() + ":" + str(part_index) + ":")
        scNode.setLengthOfSC(places[k + 1][3])
        scNodes.append(scNode)
        accNode.addChildNode(scNode, AccNode.BODY, part_index, AccNode.BEFORE)
    accNode, part_index = places[-1][0], places[-1][1]
    scNode = SC_NodeConstructor(space_charge_calculator,
                                accNode.getName() + ":" + str(part_index) + ":")
    scNode.setLengthOfSC(rest_length)
    scNodes.append(scNode)
    accNode.addChildNode(scNode, AccNode.BODY, part_index, AccNode.BEFORE)
    return scNodes
