"""orbitFinalize (reference: py/orbit/utils/orbitFinalize.py:1-12).

The reference prints the message and a stack trace, then exits the process.  Here the
condition is raised as OrbitFinalizeError so that callers (and tests) can handle it.
"""


class OrbitFinalizeError(RuntimeError):
    pass


def orbitFinalize(message=None):
    raise OrbitFinalizeError("ORBIT message: " + str(message))
