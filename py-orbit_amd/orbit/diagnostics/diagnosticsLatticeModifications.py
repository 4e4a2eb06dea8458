"""Placing diagnostics nodes in a lattice (reference:
py/orbit/diagnostics/diagnosticsLatticeModifications.py:17-124)."""
from orbit.utils import orbitFinalize
from orbit.teapot import DriftTEAPOT
from orbit.diagnostics.TeapotDiagnosticsNode import (TeapotStatLatsNodeSetMember,
                                                     TeapotMomentsNodeSetMember)

_LENGTH_TOLERANCE = 0.0001  # [m]: shorter leftover drifts are dropped


def addTeapotDiagnosticsNode(lattice, position, diagnostics_node):
    """Put diagnostics_node at position [m] of the lattice, inside drifts only: the drift
    (or drifts) covering [position, position + node length] are replaced by the node and
    the leftover drift pieces longer than 0.1 mm (:17-65)."""
    lattice.initialize()
    start, stop = position, position + diagnostics_node.getLength()
    diagnostics_node.setPosition(position)
    diagnostics_node.setLatticeLength(lattice.getLength())
    nodes = lattice.getNodes()
    i_start = i_stop = -1
    z = 0.0
    for ind, node in enumerate(nodes):
        if z <= start <= z + node.getLength():
            i_start = ind
        if z <= stop <= z + node.getLength():
            i_stop = ind
        z += node.getLength()
    for node in nodes[i_start:i_stop + 1]:
        if not isinstance(node, DriftTEAPOT):
            orbitFinalize("We have non-drift element at the place of the diagnostics! Stop! "
                          "node=" + node.getName() + " type=" + node.getType())
    pos = lattice.getNodePositionsDict()
    first, last = nodes[i_start], nodes[i_stop]
    new_nodes = [diagnostics_node]
    if start > pos[first][0] + _LENGTH_TOLERANCE:
        d = DriftTEAPOT(first.getName())
        d.setLength(start - pos[first][0])
        new_nodes.insert(0, d)
    if stop < pos[last][1] - _LENGTH_TOLERANCE:
        d = DriftTEAPOT(last.getName())
        d.setLength(pos[last][1] - stop)
        new_nodes.append(d)
    nodes[i_start:i_stop + 1] = new_nodes
    lattice.initialize()


def addTeapotDiagnosticsNodeAsChild(lattice, AccNode, diagnostics_node):
    """Attach the node at the entrance of AccNode (:67-69)."""
    AccNode.addChildNode(diagnostics_node, AccNode.ENTRANCE, 0, AccNode.BEFORE)
    lattice.initialize()


def _node_set(lattice, file_out, controller_name, make):
    controller = diagnosticsNodeSetController(file_out, controller_name)
    lattice.initialize()
    for node in lattice.getNodes():
        member = make()
        member.setPosition(lattice.getNodePositionsDict()[node][0])
        member.setLatticeLength(lattice.getLength())
        node.addChildNode(member, node.BODY)
        controller._nodelist.append(member)
    return controller


def addTeapotStatLatsNodeSet(lattice, filename):
    """A StatLats member at the start of the body of every lattice node, all writing to
    one file; returns their controller (:71-87)."""
    file_out = open(filename, "w")
    return _node_set(lattice, file_out, "StatLats Set Controller",
                     lambda: TeapotStatLatsNodeSetMember(file_out))


def addTeapotMomentsNodeSet(lattice, filename, order, nodispersion=True, emitnorm=False):
    """A Moments member at the start of the body of every lattice node, all writing to one
    file; returns their controller (:89-106)."""
    file_out = open(filename, "w")
    return _node_set(lattice, file_out, "Moment Set Controller",
                     lambda: TeapotMomentsNodeSetMember(file_out, order, nodispersion,
                                                        emitnorm, "moments"))


class diagnosticsNodeSetController:
    """Switches a node set on and off and moves it to a new file (:108-124)."""

    def __init__(self, file=None, name="Diagnostic Set Controller"):
        self._nodelist = []
        self._file = file

    def activate(self):
        for node in self._nodelist:
            node.activate()

    def deactivate(self):
        for node in self._nodelist:
            node.deactivate()

    def resetFile(self, filename):
        self._file.close()
        self._file = open(filename, "w")
        for node in self._nodelist:
            node.resetFile(self._file)
