"""teapot_base -- TEAPOT element maps on the device.

Same functions, argument order and meaning as the reference module
(src/teapot/wrap_teapotbase.cc:488-512; maps in src/teapot/teapotbase.cc).  Each call
queues the map on the bunch; queued maps are applied by one fused kernel when the bunch
is next needed, with results identical to applying them one at a time.  useCharge is
accepted and ignored, as in the reference.
"""
import ctypes as C

import _obnative as N


def _h(bunch):
    return bunch._h


def _batched(bunch):
    return getattr(bunch, "_batch", None)


def rotatexy(bunch, angle):
    if _batched(bunch):
        return bunch._batch.trackSequence([("rotatexy", angle)])
    N.call("ob_tpb_rotatexy", _h(bunch), float(angle))


def drift(bunch, length):
    if _batched(bunch):
        return bunch._batch.trackSequence([("drift", length)])
    N.call("ob_tpb_drift", _h(bunch), float(length))


def wrapbunch(bunch, length):
    if _batched(bunch):
        return bunch._batch.trackSequence([("wrapbunch", length)])
    N.call("ob_tpb_wrapbunch", _h(bunch), float(length))


def kick(bunch, kx, ky, kE, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("kick", kx, ky, kE, useCharge)])
    N.call("ob_tpb_kick", _h(bunch), float(kx), float(ky), float(kE), int(useCharge))


def multp(bunch, pole, kl, skew, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("multp", pole, kl, skew, useCharge)])
    N.call("ob_tpb_multp", _h(bunch), int(pole), float(kl), int(skew), int(useCharge))


def multpfringeIN(bunch, pole, kl, skew, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("multpfringeIN", pole, kl, skew, useCharge)])
    N.call("ob_tpb_multpfringeIN", _h(bunch), int(pole), float(kl), int(skew), int(useCharge))


def multpfringeOUT(bunch, pole, kl, skew, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("multpfringeOUT", pole, kl, skew, useCharge)])
    N.call("ob_tpb_multpfringeOUT", _h(bunch), int(pole), float(kl), int(skew), int(useCharge))


def quad1(bunch, length, kq, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("quad1", length, kq, useCharge)])
    N.call("ob_tpb_quad1", _h(bunch), float(length), float(kq), int(useCharge))


def quad2(bunch, length):
    if _batched(bunch):
        return bunch._batch.trackSequence([("quad2", length)])
    N.call("ob_tpb_quad2", _h(bunch), float(length))


def quadfringeIN(bunch, kq, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("quadfringeIN", kq, useCharge)])
    N.call("ob_tpb_quadfringeIN", _h(bunch), float(kq), int(useCharge))


def quadfringeOUT(bunch, kq, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("quadfringeOUT", kq, useCharge)])
    N.call("ob_tpb_quadfringeOUT", _h(bunch), float(kq), int(useCharge))


class _MapOp(C.Structure):
    """ob_map_op of include/orbit_mi355x.h."""
    _fields_ = [("kind", C.c_int32), ("ia", C.c_int32), ("ib", C.c_int32), ("ic", C.c_int32),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double)]


_KIND = {"rotatexy": 1, "drift": 2, "wrapbunch": 3, "kick": 4, "multp": 5,
         "multpfringeIN": 6, "multpfringeOUT": 7, "quad1": 8, "quad2": 9,
         "quadfringeIN": 10, "quadfringeOUT": 11}


def _map_op(name, *args):
    k = _KIND[name]
    o = _MapOp(kind=k, ic=1)
    if name in ("rotatexy", "drift", "wrapbunch", "quad2"):
        (o.a,) = (float(args[0]),)
    elif name == "kick":
        o.a, o.b, o.c = (float(v) for v in args[:3])
        o.ic = int(args[3]) if len(args) > 3 else 1
    elif name.startswith("multp"):
        o.ia, o.a, o.ib = int(args[0]), float(args[1]), int(args[2])
        o.ic = int(args[3]) if len(args) > 3 else 1
    elif name == "quad1":
        o.a, o.b = float(args[0]), float(args[1])
        o.ic = int(args[2]) if len(args) > 2 else 1
    else:  # quadfringeIN / quadfringeOUT
        o.b = float(args[0])
        o.ic = int(args[1]) if len(args) > 1 else 1
    return o


def trackSequence(bunch, ops):
    """Apply a list of maps, e.g. [("quadfringeIN", kq), ("quad1", 0.25, kq), ("drift", 0.5)],
    in one call and one fused launch (no reference counterpart; same results as calling the
    functions above in order)."""
    arr = (_MapOp * len(ops))(*[_map_op(op[0], *op[1:]) for op in ops])
    N.call("ob_track_sequence", _h(bunch), arr, len(ops))
