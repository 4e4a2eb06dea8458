"""Histogram of one bunch coordinate (reference: py/orbit/diagnostics/profiles.py:14-91).

The extrema and the binning both run on the device (BunchExtremaCalculator,
Grid1D.binBunch); only the histogram comes back to be written."""
import os

from spacecharge import Grid1D
from orbit_utils import BunchExtremaCalculator
from orbit.utils import orbitFinalize

_AXES = {"x": 0, "px": 1, "y": 2, "py": 3, "z": 4, "dE": 5}


def profiles(bunch, coord, histogram, steps=100, Min=1.0, Max=-1.0):
    """Bin coordinate coord ("x", "px", "y", "py", "z" or "dE") of the bunch into `steps`
    cells over [Min, Max] (the bunch's extent when Min >= Max) with linear weights and
    macrosizes, write the histogram to the file `histogram` and return it as a Grid1D."""
    grid1D = Grid1D(steps)
    comm = bunch.getMPIComm()
    nParts = bunch.getSizeGlobal()
    if nParts < 2:
        with open(histogram, "w") as f:
            f.write("bunch has only number of particles =" + str(nParts) + os.linesep)
        return grid1D
    axis = _AXES.get(coord, -1)
    if axis < 0:
        orbitFinalize("Error: profiles(bunch,coord,...)" + os.linesep +
                      "coord is not x,xp,y,yp,z, or dE." + os.linesep + "Stop." + os.linesep)
        return None
    if Min >= Max:
        calc = BunchExtremaCalculator()
        xyz = calc.extremaXYZ(bunch)
        pde = calc.extremaXpYpdE(bunch)
        lims = [xyz[0:2], pde[0:2], xyz[2:4], pde[2:4], xyz[4:6], pde[4:6]]
        Min, Max = lims[axis]
    if Min == Max:
        with open(histogram, "w") as f:
            f.write("No histogram min=" + str(Min) + " max=" + str(Max) + os.linesep)
        return grid1D
    grid1D.setGridZ(Min, Max)
    grid1D.binBunch(bunch, axis)
    grid1D.synchronizeMPI(comm)
    total = grid1D.getSum()
    with open(histogram, "w") as f:
        f.write("% " + " Min, Max = %15.8g , %15.8g " % (Min, Max) +
                " nSteps = %5d " % steps + " hist_sum = %15.8g" % total + os.linesep)
        for ind in range(steps):
            f.write(" %15.8g  %15.8g " % (grid1D.getGridZ(ind), grid1D.getValueOnGrid(ind)) +
                    os.linesep)
    return grid1D
