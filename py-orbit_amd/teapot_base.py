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


def quad3(bunch, length, kq, useCharge=1):
    """quad3 (teapotbase.cc:761-764) does nothing."""
    N.call("ob_tpb_quad3", _h(bunch), float(length), float(kq), int(useCharge))


def drifti(bunch, i, length):
    """drifti (teapotbase.cc:120-141): one particle, through the host copy."""
    if _batched(bunch):
        raise NotImplementedError("drifti on a bunch of a batch")
    N.call("ob_tpb_drifti", _h(bunch), int(i), float(length))


def wedgerotate(bunch, e, frinout):
    if _batched(bunch):
        return bunch._batch.trackSequence([("wedgerotate", e, frinout)])
    N.call("ob_tpb_wedgerotate", _h(bunch), float(e), int(frinout))


def wedgedrift(bunch, e, inout):
    if _batched(bunch):
        return bunch._batch.trackSequence([("wedgedrift", e, inout)])
    N.call("ob_tpb_wedgedrift", _h(bunch), float(e), int(inout))


def wedgebend(bunch, e, inout, rho, nsteps):
    if _batched(bunch):
        return bunch._batch.trackSequence([("wedgebend", e, inout, rho, nsteps)])
    N.call("ob_tpb_wedgebend", _h(bunch), float(e), int(inout), float(rho), int(nsteps))


def _mult_arrays(vecnum, poles, kls, skews):
    """The three sequences of wedgebendCF as C arrays (wrap_teapotbase.cc:420-455 checks:
    sequences of length vecnum)."""
    if len(poles) != vecnum or len(kls) != vecnum or len(skews) != vecnum:
        raise ValueError("teapotbase - wedgebendCF - size of sequences with poles, kls or "
                         "skews are wrong!")
    return ((C.c_int * max(vecnum, 1))(*[int(v) for v in poles]),
            (C.c_double * max(vecnum, 1))(*[float(v) for v in kls]),
            (C.c_int * max(vecnum, 1))(*[int(v) for v in skews]))


def wedgebendCF(bunch, e, inout, rho, vecnum, poles, kls, skews, nsteps, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("wedgebendCF", e, inout, rho, vecnum, poles, kls,
                                            skews, nsteps, useCharge)])
    pa, ka, sa = _mult_arrays(vecnum, poles, kls, skews)
    N.call("ob_tpb_wedgebendCF", _h(bunch), float(e), int(inout), float(rho), int(vecnum),
           pa, ka, sa, int(nsteps), int(useCharge))


def bend1(bunch, length, th):
    if _batched(bunch):
        return bunch._batch.trackSequence([("bend1", length, th)])
    N.call("ob_tpb_bend1", _h(bunch), float(length), float(th))


def bend2(bunch, length):
    if _batched(bunch):
        return bunch._batch.trackSequence([("bend2", length)])
    N.call("ob_tpb_bend2", _h(bunch), float(length))


def bend3(bunch, th):
    if _batched(bunch):
        return bunch._batch.trackSequence([("bend3", th)])
    N.call("ob_tpb_bend3", _h(bunch), float(th))


def bend4(bunch, th):
    if _batched(bunch):
        return bunch._batch.trackSequence([("bend4", th)])
    N.call("ob_tpb_bend4", _h(bunch), float(th))


def bendfringeIN(bunch, rho):
    if _batched(bunch):
        return bunch._batch.trackSequence([("bendfringeIN", rho)])
    N.call("ob_tpb_bendfringeIN", _h(bunch), float(rho))


def bendfringeOUT(bunch, rho):
    if _batched(bunch):
        return bunch._batch.trackSequence([("bendfringeOUT", rho)])
    N.call("ob_tpb_bendfringeOUT", _h(bunch), float(rho))


def soln(bunch, length, B, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("soln", length, B, useCharge)])
    N.call("ob_tpb_soln", _h(bunch), float(length), float(B), int(useCharge))


def RingRF(bunch, ring_length, harmonic_numb, voltage, phase_s, useCharge=1):
    if _batched(bunch):
        return bunch._batch.trackSequence([("RingRF", ring_length, harmonic_numb, voltage,
                                            phase_s, useCharge)])
    N.call("ob_tpb_RingRF", _h(bunch), float(ring_length), int(harmonic_numb), float(voltage),
           float(phase_s), int(useCharge))


class _MapOp(C.Structure):
    """ob_map_op of include/orbit_mi355x.h."""
    _fields_ = [("kind", C.c_int32), ("ia", C.c_int32), ("ib", C.c_int32), ("ic", C.c_int32),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("n", C.c_int32), ("pad", C.c_int32), ("poles", C.POINTER(C.c_int)),
                ("kls", C.POINTER(C.c_double)), ("skews", C.POINTER(C.c_int))]


_KIND = {"rotatexy": 1, "drift": 2, "wrapbunch": 3, "kick": 4, "multp": 5,
         "multpfringeIN": 6, "multpfringeOUT": 7, "quad1": 8, "quad2": 9,
         "quadfringeIN": 10, "quadfringeOUT": 11, "bend1": 12, "bend2": 13, "bend3": 14,
         "bend4": 15, "bendfringeIN": 16, "bendfringeOUT": 17, "wedgerotate": 18,
         "wedgedrift": 19, "wedgebend": 20, "wedgebendCF": 21, "soln": 22, "RingRF": 23}


def _map_op(name, *args, keep=None):
    """One ob_map_op; C arrays it points to are appended to `keep` (the caller holds them
    until the call returns)."""
    k = _KIND[name]
    o = _MapOp(kind=k, ic=1)
    if name in ("rotatexy", "drift", "wrapbunch", "quad2", "bend2", "bend3", "bend4",
                "bendfringeIN", "bendfringeOUT"):
        (o.a,) = (float(args[0]),)
    elif name == "bend1":
        o.a, o.b = float(args[0]), float(args[1])
    elif name in ("wedgerotate", "wedgedrift"):
        o.a, o.ia = float(args[0]), int(args[1])
    elif name == "wedgebend":
        o.a, o.ia, o.b, o.ib = float(args[0]), int(args[1]), float(args[2]), int(args[3])
    elif name == "wedgebendCF":
        e, inout, rho, vecnum, poles, kls, skews, nsteps = args[:8]
        o.a, o.ia, o.b, o.ib, o.n = float(e), int(inout), float(rho), int(nsteps), int(vecnum)
        o.ic = int(args[8]) if len(args) > 8 else 1
        arrs = _mult_arrays(int(vecnum), poles, kls, skews)
        keep.append(arrs)
        o.poles, o.kls, o.skews = arrs
    elif name == "soln":
        o.a, o.b = float(args[0]), float(args[1])
        o.ic = int(args[2]) if len(args) > 2 else 1
    elif name == "RingRF":
        o.a, o.ia, o.b, o.c = float(args[0]), int(args[1]), float(args[2]), float(args[3])
        o.ic = int(args[4]) if len(args) > 4 else 1
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
    keep = []
    arr = (_MapOp * len(ops))(*[_map_op(op[0], *op[1:], keep=keep) for op in ops])
    N.call("ob_track_sequence", _h(bunch), arr, len(ops))
