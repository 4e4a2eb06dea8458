"""Accelerator lattice tree and its traversal (reference: py/orbit/lattice).

A node has ENTRANCE children, nParts body parts each with BEFORE and AFTER children,
and EXIT children.  trackActions visits them in that order, running the actions
container's ENTRANCE / BODY / EXIT actions on the node itself
(AccNode.py:302-350, AccLattice.py:243-255).
"""
import os

from orbit.utils import NamedObject, ParamsDictObject, TypedObject, orbitFinalize


class AccActionsContainer(NamedObject):
    """Lists of callables run at a node's entrance, body parts and exit
    (AccActionsContainer.py)."""

    ENTRANCE = 0
    BODY = 1
    EXIT = 2
    BEFORE = 0
    AFTER = 1

    def __init__(self, name="no name container"):
        NamedObject.__init__(self, name)
        self.__actions = ([], [], [])

    def addAction(self, action, place):
        self.__actions[place].append(action)

    def removeAction(self, action, place):
        self.__actions[place].remove(action)

    def getActions(self, place):
        return self.__actions[place]

    def performActions(self, paramsDict, place):
        for action in self.__actions[place]:
            action(paramsDict)


class AccNode(NamedObject, TypedObject, ParamsDictObject):
    ENTRANCE = AccActionsContainer.ENTRANCE
    BODY = AccActionsContainer.BODY
    EXIT = AccActionsContainer.EXIT
    BEFORE = AccActionsContainer.BEFORE
    AFTER = AccActionsContainer.AFTER

    def __init__(self, name="no name", type_in="generic"):
        NamedObject.__init__(self, name)
        TypedObject.__init__(self, type_in)
        ParamsDictObject.__init__(self)
        self.__nParts = 1
        self.__length = 0.0
        self.__partLengths = [0.0]
        self.__activePart = 0
        self.__entrance = []
        self.__body = [[[], []]]
        self.__exit = []
        self._setPartsLengthEvenly(1)

    # ---- parts and lengths --------------------------------------------------
    def setnParts(self, n=1):
        if self.getNumberOfBodyChildren() != 0:
            orbitFinalize("You cannot set the number of AccNode parts after you added "
                          "children! Name of element=" + self.getName())
        self._setPartsLengthEvenly(n)
        self.initialize()

    def getnParts(self):
        return self.__nParts

    def _setPartsLengthEvenly(self, n=1):
        if self.getNumberOfBodyChildren() != 0:
            orbitFinalize("AccNode._setPartsLengthEvenly would remove the existing body "
                          "children of " + self.getName())
        self.__nParts = n
        self.__partLengths = [self.__length / n for _ in range(n)]
        self.__body = [[[], []] for _ in range(n)]

    def setLength(self, L=0.0, index=-1):
        L = float(L)
        if abs(L) < 1.0e-36:
            L = 0.0
        if index >= 0:
            self.__partLengths[index] = L
            return
        self.__length = L
        self._setPartsLengthEvenly(self.__nParts)
        self.initialize()

    def getLength(self, index=-1):
        if index >= 0:
            return self.__partLengths[index]
        return self.__length

    def getActivePartIndex(self):
        return self.__activePart

    def setActivePartIndex(self, index):
        self.__activePart = index

    def initialize(self):
        """Hook for subclasses whose parts are not of equal length."""
        pass

    # ---- children -------------------------------------------------------------
    def _place_list(self, place, part_index=0, place_in_part=AccActionsContainer.BEFORE):
        if place == AccNode.ENTRANCE:
            return self.__entrance
        if place == AccNode.EXIT:
            return self.__exit
        if place != AccNode.BODY:
            orbitFinalize("AccNode.addChildNode: place must be ENTRANCE, BODY or EXIT; got "
                          + str(place))
        return self.__body[part_index][place_in_part]

    def addChildNode(self, node, place, part_index=0, place_in_part=AccActionsContainer.BEFORE):
        self._place_list(place, part_index, place_in_part).append(node)

    def getChildNodes(self, place, part_index=0, place_in_part=AccActionsContainer.BEFORE):
        return self._place_list(place, part_index, place_in_part)

    def getNumberOfBodyChildren(self):
        body = getattr(self, "_AccNode__body", None)
        if body is None:
            return 0
        return sum(len(b) + len(a) for b, a in body)

    def getNumberOfChildren(self):
        return len(self.__entrance) + len(self.__exit) + self.getNumberOfBodyChildren()

    def getBodyChildren(self):
        nodes = []
        for before, after in self.__body:
            nodes += before
            nodes += after
        return nodes

    def getAllChildren(self):
        return list(self.__entrance) + self.getBodyChildren() + list(self.__exit)

    # ---- traversal ------------------------------------------------------------
    def trackActions(self, actionsContainer, paramsDict={}):
        paramsDict["node"] = self
        parentNode = paramsDict.get("parentNode", None)
        if "path_length" not in paramsDict:
            paramsDict["path_length"] = 0.0
        has_length = self.getLength() > 0.0
        self.__activePart = -1
        actionsContainer.performActions(paramsDict, AccNode.ENTRANCE)
        for node in self.__entrance:
            paramsDict["node"] = node
            paramsDict["parentNode"] = self
            node.trackActions(actionsContainer, paramsDict)
        for i in range(self.__nParts):
            paramsDict["node"] = self
            paramsDict["parentNode"] = parentNode
            self.__activePart = i
            for node in self.__body[i][AccNode.BEFORE]:
                paramsDict["parentNode"] = self
                node.trackActions(actionsContainer, paramsDict)
            paramsDict["node"] = self
            paramsDict["parentNode"] = parentNode
            actionsContainer.performActions(paramsDict, AccNode.BODY)
            if has_length:
                paramsDict["path_length"] += self.getLength(i)
            for node in self.__body[i][AccNode.AFTER]:
                paramsDict["parentNode"] = self
                node.trackActions(actionsContainer, paramsDict)
        self.__activePart = -1
        for node in self.__exit:
            paramsDict["node"] = node
            paramsDict["parentNode"] = self
            node.trackActions(actionsContainer, paramsDict)
        paramsDict["node"] = self
        paramsDict["parentNode"] = parentNode
        actionsContainer.performActions(paramsDict, AccNode.EXIT)

    def structureToText(self, txt="", txt_shift=""):
        sp = " "
        txt += txt_shift + "==== START AccNode = " + self.getName() + " L=" + str(self.getLength()) + os.linesep
        for node in self.__entrance:
            txt = node.structureToText(txt, txt_shift + sp * 2)
        for i in range(self.__nParts):
            txt += txt_shift + sp * 3 + "==== BODY Part ind.=" + str(i) + " L=" + str(self.getLength(i)) + os.linesep
            for node in self.__body[i][AccNode.BEFORE] + self.__body[i][AccNode.AFTER]:
                txt = node.structureToText(txt, txt_shift + sp * 5)
        for node in self.__exit:
            txt = node.structureToText(txt, txt_shift + sp * 2)
        txt += txt_shift + "==== END of AccNode = " + self.getName() + os.linesep
        return txt


class AccNodeBunchTracker(AccNode):
    """Base of nodes that track a bunch (AccNodeBunchTracker.py:14-27)."""

    def __init__(self, name="no name", type_in="bunch tracker"):
        AccNode.__init__(self, name, type_in)

    def trackBunch(self, bunch, paramsDict={}, actionContainer=None):
        if actionContainer is None:
            actionContainer = AccActionsContainer("Bunch Tracking")
        paramsDict["bunch"] = bunch

        def track(params):
            params["node"].track(params)

        actionContainer.addAction(track, AccActionsContainer.BODY)
        self.trackActions(actionContainer, paramsDict)
        actionContainer.removeAction(track, AccActionsContainer.BODY)

    def track(self, paramsDict):
        pass


class AccLattice(NamedObject, TypedObject):
    """Ordered first-level nodes (AccLattice.py)."""

    ENTRANCE = AccActionsContainer.ENTRANCE
    BODY = AccActionsContainer.BODY
    EXIT = AccActionsContainer.EXIT
    BEFORE = AccActionsContainer.BEFORE
    AFTER = AccActionsContainer.AFTER

    def __init__(self, name="no name"):
        NamedObject.__init__(self, name)
        TypedObject.__init__(self, "lattice")
        self.__length = 0.0
        self.__initialized = False
        self.__children = []
        self.__positions = {}

    def initialize(self):
        """Initialize every node and compute node positions (AccLattice.py:36-80)."""
        seen = set()
        for node in self.__children:
            if id(node) in seen:
                orbitFinalize("The AccLattice class instance should not have duplicate nodes! "
                              "Name of node=" + node.getName())
            seen.add(id(node))
            node.initialize()
        pos = {}
        run = [0.0]

        def on_exit(params):
            node = params["node"]
            if isinstance(params["parentNode"], AccLattice):
                start = run[0]
                run[0] += node.getLength()
                pos[node] = (start, run[0])

        actions = AccActionsContainer()
        actions.addAction(on_exit, AccNode.EXIT)
        self.trackActions(actions, {})
        self.__length = run[0]
        self.__positions = pos
        self.__initialized = True

    def isInitialized(self):
        return self.__initialized

    def addNode(self, node, index=-1):
        if isinstance(node, AccNode):
            if index < 0:
                self.__children.append(node)
            else:
                self.__children.insert(index, node)
            self.__initialized = False

    def getNodes(self):
        return self.__children

    def setNodes(self, nodes):
        self.__children = nodes

    def getNodeIndex(self, node):
        return self.__children.index(node)

    def getNodeForName(self, name):
        nodes = self.getNodesForName(name)
        if len(nodes) == 1:
            return nodes[0]
        if not nodes:
            return None
        orbitFinalize("AccLattice.getNodeForName found many nodes for name=" + name)

    def getNodesForName(self, name):
        return [n for n in self.__children if n.getName().find(name) == 0]

    def getNodesOfClass(self, cls):
        return [n for n in self.__children if isinstance(n, cls)]

    def getNodePositionsDict(self):
        return self.__positions

    def getLength(self):
        return self.__length

    def trackActions(self, actionsContainer, paramsDict={}, index_start=-1, index_stop=-1):
        """Track through nodes index_start..index_stop inclusive (AccLattice.py:243-255)."""
        paramsDict["lattice"] = self
        paramsDict["actions"] = actionsContainer
        if "path_length" not in paramsDict:
            paramsDict["path_length"] = 0.0
        if index_start < 0:
            index_start = 0
        if index_stop < 0:
            index_stop = len(self.__children) - 1
        for node in self.__children[index_start:index_stop + 1]:
            paramsDict["node"] = node
            paramsDict["parentNode"] = self
            node.trackActions(actionsContainer, paramsDict)

    def structureToText(self):
        txt = "==== START ==== Lattice =" + self.getName() + "  L=" + str(self.getLength()) + os.linesep
        for node in self.__children:
            txt += node.structureToText("")
        txt += "==== STOP  ==== Lattice =" + self.getName() + "  L=" + str(self.getLength()) + os.linesep
        return txt


__all__ = ["AccActionsContainer", "AccNode", "AccLattice", "AccNodeBunchTracker"]
