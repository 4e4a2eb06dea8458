"""teapot_base -- TEAPOT element maps on the device.

Same functions, argument order and meaning as the reference module
(src/teapot/wrap_teapotbase.cc:488-512; maps in src/teapot/teapotbase.cc).  Each call
queues the map on the bunch; queued maps are applied by one fused kernel when the bunch
is next needed, with results identical to applying them one at a time.  useCharge is
accepted and ignored, as in the reference.
"""
import _obnative as N


def _h(bunch):
    return bunch._h


def rotatexy(bunch, angle):
    N.call("ob_tpb_rotatexy", _h(bunch), float(angle))


def drift(bunch, length):
    N.call("ob_tpb_drift", _h(bunch), float(length))


def wrapbunch(bunch, length):
    N.call("ob_tpb_wrapbunch", _h(bunch), float(length))


def kick(bunch, kx, ky, kE, useCharge=1):
    N.call("ob_tpb_kick", _h(bunch), float(kx), float(ky), float(kE), int(useCharge))


def multp(bunch, pole, kl, skew, useCharge=1):
    N.call("ob_tpb_multp", _h(bunch), int(pole), float(kl), int(skew), int(useCharge))


def multpfringeIN(bunch, pole, kl, skew, useCharge=1):
    N.call("ob_tpb_multpfringeIN", _h(bunch), int(pole), float(kl), int(skew), int(useCharge))


def multpfringeOUT(bunch, pole, kl, skew, useCharge=1):
    N.call("ob_tpb_multpfringeOUT", _h(bunch), int(pole), float(kl), int(skew), int(useCharge))


def quad1(bunch, length, kq, useCharge=1):
    N.call("ob_tpb_quad1", _h(bunch), float(length), float(kq), int(useCharge))


def quad2(bunch, length):
    N.call("ob_tpb_quad2", _h(bunch), float(length))


def quadfringeIN(bunch, kq, useCharge=1):
    N.call("ob_tpb_quadfringeIN", _h(bunch), float(kq), int(useCharge))


def quadfringeOUT(bunch, kq, useCharge=1):
    N.call("ob_tpb_quadfringeOUT", _h(bunch), float(kq), int(useCharge))
