"""NamedObject, TypedObject, ParamsDictObject (reference: py/orbit/utils/*.py)."""
import os

from orbit.utils.orbitFinalize import orbitFinalize


class NamedObject:
    def __init__(self, name="no name"):
        self.__name = name

    def setName(self, name="no name"):
        self.__name = name

    def getName(self):
        return self.__name


class TypedObject:
    def __init__(self, type_in="no type"):
        self.__type = type_in

    def setType(self, type_in="no type"):
        self.__type = type_in

    def getType(self):
        return self.__type


class ParamsDictObject:
    """An object carrying a parameter dictionary (ParamsDictObject.py:12-84)."""

    def __init__(self):
        self.__params = {}

    def addParam(self, key, value):
        self.__params[key] = value

    def setParam(self, key, value):
        self.__params[key] = value

    def removeParam(self, key):
        self.__params.pop(key, None)

    def setParamsDict(self, params):
        self.__params = params

    def updateParamsDict(self, params):
        self.__params.update(params)

    def getParamsDict(self):
        return self.__params

    def hasParam(self, key):
        return key in self.__params

    def keys(self):
        return list(self.__params.keys())

    def getParam(self, key):
        if key not in self.__params:
            who = ""
            if isinstance(self, NamedObject):
                who += "Name of element = " + self.getName() + os.linesep
            if isinstance(self, TypedObject):
                who += "Type of element = " + self.getType() + os.linesep
            orbitFinalize("The object does not have a parameter for the key you requested!"
                          + os.linesep + who + "key = " + str(key))
        return self.__params[key]
