"""Shared test utilities: lattice tracing and oracle replay."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


class TraceTPB:
    """Stands in for teapot_base: records each map call."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def rec(bunch, *args):
            self.calls.append((name,) + tuple(args))
        return rec


class TraceCalc:
    """Stands in for a space-charge calculator: records trackBunch calls."""

    def __init__(self, calls, kind):
        self.calls = calls
        self.kind = kind

    def trackBunch(self, bunch, *args):
        self.calls.append((self.kind,) + tuple(args))


def trace_lattice(monkeypatch, build, kind, *sc_args, turns=1):
    """Build a lattice with traced maps and SC calculator; return the ordered calls.
    build(calc) -> lattice (with SC nodes inserted)."""
    import orbit.teapot.teapot as tt
    tpb = TraceTPB()
    monkeypatch.setattr(tt, "TPB", tpb)
    calc = TraceCalc(tpb.calls, kind)
    lat = build(calc)

    class FakeBunch:
        def hasBunchAttrInt(self, name):
            return 0

    for _ in range(turns):
        lat.trackBunch(FakeBunch())
    return tpb.calls


def replay_oracle(calls, obunch, calc):
    """Apply a recorded call list to an oracle bunch."""
    for c in calls:
        name, args = c[0], c[1:]
        if name in ("sc2p5drb", "sc2p5d", "sc3d", "slice2d"):
            calc.trackBunch(obunch, *args)
        else:
            if name in ("quad1", "kick", "multp", "multpfringeIN", "multpfringeOUT", "soln",
                        "quadfringeIN", "quadfringeOUT"):
                # drop useCharge (ignored by the reference maps)
                nargs = {"quad1": 2, "kick": 3, "multp": 3, "multpfringeIN": 3, "soln": 2,
                         "multpfringeOUT": 3, "quadfringeIN": 1, "quadfringeOUT": 1}[name]
                args = args[:nargs]
            obunch.apply(name, *args)
