"""BunchExtremaCalculator under the reference's module name `orbit_utils`
(src/utils/bunch/BunchExtremaCalculator.cc:37-208, wrap_bunch_extrema_calculator.cc:44-109).

The extrema are found on the device (ob_bunch_extrema) after the bunch's queued maps;
the coordinates are not downloaded.  An empty bunch gives (DBL_MAX, -DBL_MAX) pairs, as in
the reference."""
import numpy as np

import _obnative as N


class BunchExtremaCalculator:
    def _all(self, bunch):
        if not hasattr(bunch, "_h"):
            raise TypeError("BunchExtremaCalculator: the input parameter is not a Bunch")
        out = np.zeros(12)
        N.call("ob_bunch_extrema", bunch._h, N.dptr(out))
        return out

    def extremaXYZ(self, bunch):
        """(xMin, xMax, yMin, yMax, zMin, zMax)."""
        e = self._all(bunch)
        return tuple(float(e[k]) for k in (0, 1, 4, 5, 8, 9))

    def extremaXpYpdE(self, bunch):
        """(xpMin, xpMax, ypMin, ypMax, dE_Min, dE_Max)."""
        e = self._all(bunch)
        return tuple(float(e[k]) for k in (2, 3, 6, 7, 10, 11))

    def extremaZ(self, bunch):
        """(zMin, zMax)."""
        e = self._all(bunch)
        return float(e[8]), float(e[9])

    def extremaR(self, bunch):
        """Largest distance of a particle from the unweighted mean (x, y)."""
        if not hasattr(bunch, "_h"):
            raise TypeError("BunchExtremaCalculator: the input parameter is not a Bunch")
        out = np.zeros(1)
        N.call("ob_bunch_extrema_r", bunch._h, N.dptr(out))
        return float(out[0])
