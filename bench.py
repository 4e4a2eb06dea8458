"""Benchmark: particle-pushes/s of TEAPOT tracking with the 3D space-charge kick on one
MI355X.

Workload (default, `ring3d`): 4,000,000 macroparticles (4D waterbag, radii 6 mm / 1 mrad,
z Gaussian sigma 0.3 m, dE sigma 1 MeV, protons at 1 GeV) tracked through a 24-cell ring
of FODO cells with two sextupoles each (k1 = +-0.5 m^-2, k2 = +-2 m^-3), a
SpaceChargeCalc3D(256, 256, 128) node every ~0.5 m.  One particle-push is one
macroparticle advanced through one space-charge interval (the maps between two SC nodes
plus the SC kick with its extrema, deposit, Poisson solve and kick).  One bench step is
one ring cell (its maps and its SC nodes); `value` = N x SC intervals / wall time, with
the bunch already resident in HBM.

`fodo2p5d` runs the 2.5D workload instead (100k particles, 128x128x64 SpaceChargeCalc2p5Drb,
FODO cells, 12 SC nodes per cell).  `batch64` tracks 64 independent bunches together (sizes
uniform in 50k-2M, Gaussian or KV, emittance 1-20 mm mrad, intensity 1e12-1e14, 2.5D grids
64^2/128^2/256^2 x 64, FODO cells): one cell is captured in a hipGraph and replayed, one
bench step per replay.

The JSON line carries `roofline` (dominant kernel: algorithmic bytes per launch / its
mean HIP-event duration over a profiled pass) and `cpu_baseline` (the CPU oracle on a
bounded sample of the same workload, timed on this host).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "py-orbit_amd")]

HBM_PEAK_GBS = 8000.0
METRIC = "particle-pushes/s incl. SC kick, 1 MI355X; field/\u03b5-growth err vs high-N FP64"
# rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the same workload (bytes per dispatch, already
# scaled by the copy-kernel yardstick); see the commands recorded in the file.
TRAFFIC_FILES = {"ring3d": "profiles/ring3d_hbm_traffic.md",
                 "fodo2p5d": "profiles/fodo2p5d_hbm_traffic.md"}
# profiler label -> kernel name (prefix) in the rocprofv3 summaries
ROCPROF_NAME = {"fft3_strided_conv": "fft3_strided<2,", "fft3_strided_fwd": "fft3_strided<0,",
                "fft3_strided_inv": "fft3_strided<1,", "push_extrema": "push_kernel<true,",
                "push": "push_kernel<false,", "fft_cols": "fft_cols<false,",
                "fft3_z_r2c": "fft3_z_r2c<", "fft3_z_c2r": "fft3_z_c2r<"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ring3d",
                    choices=["ring3d", "fodo2p5d", "fodo2p5d_graph", "batch64", "denoise",
                             "batch_denoise"])
    ap.add_argument("--pairs", type=int, default=20000, help="denoise: training pairs")
    ap.add_argument("--bunches", type=int, default=32, help="batch_denoise: bunches")
    ap.add_argument("--dn-pairs", type=int, default=4096,
                    help="batch_denoise: generated training pairs")
    ap.add_argument("--ref-bunches", type=int, default=2,
                    help="batch_denoise: bunches re-tracked alone (plain, n and 16M particles)")
    ap.add_argument("--epochs", type=int, default=2, help="denoise: training epochs")
    ap.add_argument("--dn-scales", type=int, default=5, help="denoise: loss pyramid scales")
    ap.add_argument("--turns", type=int, default=5, help="denoise: ring turns of the tracking test")
    ap.add_argument("--capture-turns", type=int, default=5,
                    help="denoise: ring turns whose SC nodes add training pairs")
    ap.add_argument("--n", type=int, default=0, help="override particle count")
    ap.add_argument("--intensity", type=float, default=0.0,
                    help="protons per bunch (default: the workload's)")
    ap.add_argument("--lds-budget", type=int, default=0,
                    help="ring3d: LDS bytes per deposit/kick tile (0: the calculator's default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="ring3d: skip the 4N reference run (emittance growth, field error)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile-steps", type=int, default=2)
    return ap.parse_args()


def build_workload(args):
    import spacecharge
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from orbit.space_charge.sc3d import setSC3DAccNodes
    from workloads import (INTENSITY, PIPE_RADIUS, fodo_lattice, gaussian_coords, make_bunch,
                           ring_lattice, waterbag_coords)
    if args.workload == "ring3d":
        n = args.n or 4_000_000
        grid = (256, 256, 128)
        coords = waterbag_coords(n)
        lat = ring_lattice(24)
        calc = spacecharge.SpaceChargeCalc3D(*grid)
        if args.lds_budget:
            calc.setDepositMode(True, args.lds_budget)
        scn = setSC3DAccNodes(lat, 0.49, calc)
        cell_len = 6  # QF S1 D1 QD S2 D2 (the ring's turn counter node follows the cells)
    else:
        n = args.n or 100_000
        grid = (128, 128, 64)
        coords = gaussian_coords(n)
        lat = fodo_lattice(100)
        calc = spacecharge.SpaceChargeCalc2p5Drb(*grid)
        scn = setSC2p5DrbAccNodes(lat, 0.49, calc, PIPE_RADIUS)
        cell_len = 4
    b = make_bunch(coords, intensity=args.intensity or INTENSITY)
    # SC nodes per cell (children of the cell's first-level nodes)
    nodes = lat.getNodes()
    cells = []
    for i in range(0, len(nodes) - (len(nodes) % cell_len), cell_len):
        ids = {id(x) for x in scn}
        nsc = sum(1 for nd in nodes[i:i + cell_len] for ch in nd.getBodyChildren() if id(ch) in ids)
        cells.append((i, i + cell_len - 1, nsc))
    return b, lat, calc, cells, n, grid, coords


def beam_stats(b):
    """Beam quality after the run, from device reductions without downloading the bunch:
    rms emittances (BunchTwissAnalysis) and the fraction of particles beyond 4 rms sizes
    from the mean in x and in y (a one-cell Grid1D over mean +- 4 sigma counts the
    particles inside)."""
    import spacecharge
    from bunch import BunchTwissAnalysis
    tw = BunchTwissAnalysis()
    tw.analyzeBunch(b)
    n = b.getSize()
    out = {"emit_x": tw.getEmittance(0), "emit_y": tw.getEmittance(1)}
    for name, ax in (("x", 0), ("y", 2)):
        mu = tw.getAverage(ax)
        sig = math.sqrt(max(tw.getCorrelation(ax, ax), 0.0))
        g = spacecharge.Grid1D(1, mu - 4 * sig, mu + 4 * sig)
        g.binBunchByParticle(b, ax)
        out["halo_4sigma_" + name] = (n - g.getSum()) / n if n else 0.0
    return out


def run_cells(lat, b, cells, k0, count):
    pushes_sc = 0
    for j in range(count):
        i0, i1, nsc = cells[(k0 + j) % len(cells)]
        lat.trackBunch(b, index_start=i0, index_stop=i1)
        pushes_sc += nsc
    return pushes_sc


def kernel_bytes(workload, n, grid):
    """Algorithmic (compulsory) HBM bytes per launch of each kernel."""
    nx, ny, nz = grid

    def pad(v):
        m = 1
        while m < 2 * v - 1:
            m <<= 1
        return max(m, 2)
    M0, M1, M2 = pad(nx), pad(ny), pad(nz)
    b = {"push_extrema": 96 * n, "push": 96 * n}
    if workload == "ring3d":
        ng = nx * ny * nz
        H2 = M2 // 2 + 1
        qx, qy = M0 // 2 + 1, M1 // 2 + 1
        b.update({
            "deposit3d": 24 * n + 8 * ng,
            "fill_zero": 8 * ng,
            "fft3_z_r2c": 8 * ng + 16 * nx * ny * H2,
            "fft3_strided_fwd": 16 * nx * ny * H2 + 16 * nx * M1 * H2,
            "fft3_strided_conv": 32 * nx * M1 * H2 + 8 * qx * qy * H2,
            "fft3_strided_inv": 16 * nx * M1 * H2 + 16 * nx * ny * H2,
            "fft3_z_c2r": 16 * nx * ny * H2 + 8 * ng,
        })
    else:
        H = M1 // 2 + 1
        b.update({
            "sc2d_bin": 24 * n + 8 * nx * ny,
            "sc2d_scatter": 32 * n,
            "sc2d_tdeposit": 16 * n + 8 * nx * ny,
            "fft_rows_r2c": 8 * nx * ny + 16 * nx * H,
            "fft_cols": 32 * nx * H + 16 * M0 * H,
            "fft_rows_c2r": 16 * nx * H + 8 * nx * ny,
        })
    return b


def measured_traffic(workload, label):
    """HBM bytes per launch of `label` from the committed rocprofv3 summary, or None."""
    path = os.path.join(ROOT, TRAFFIC_FILES[workload])
    name = ROCPROF_NAME.get(label, label)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == "traffic_bytes" and f[1].startswith(name):
            return int(f[2]) + int(f[3])
    return None


def profile(N, lat, b, cells, steps):
    import ctypes as C
    N.call("ob_profile_enable", 1)
    run_cells(lat, b, cells, 0, steps)
    N.synchronize()
    buf = C.create_string_buffer(1 << 16)
    N.call("ob_profile_report", buf, len(buf))
    N.call("ob_profile_enable", 0)
    rows = {}
    for line in buf.value.decode().strip().splitlines():
        name, cnt, tot, mn, mx = line.split()
        rows[name] = (int(cnt), float(tot), float(mn), float(mx))
    return rows


# teapot_base maps whose trailing useCharge argument the reference ignores (the oracle
# takes the arguments without it)
_ORACLE_NARGS = {"quad1": 2, "kick": 3, "multp": 3, "multpfringeIN": 3, "multpfringeOUT": 3,
                 "quadfringeIN": 1, "quadfringeOUT": 1, "soln": 2}


def lattice_program(workload):
    """The workload's SC intervals as the lattice layer issues them: a list of
    (maps, sc_args), maps being the teapot_base calls (name, *args) in order and sc_args the
    calculator's trackBunch arguments.  Recorded once by running the lattice with stand-ins
    for teapot_base and the calculator (no device work)."""
    import orbit.teapot.teapot as tt
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from orbit.space_charge.sc3d import setSC3DAccNodes
    from workloads import PIPE_RADIUS, fodo_lattice, ring_lattice
    calls = []

    class Maps:
        def __getattr__(self, name):
            return lambda bunch, *a: calls.append((name,) + a)

    class Calc:
        def trackBunch(self, bunch, *a):
            calls.append(("SC",) + a)

    class Stand:
        def hasBunchAttrInt(self, name):
            return 0

    saved = tt.TPB
    tt.TPB = Maps()
    try:
        if workload == "ring3d":
            lat = ring_lattice(24)
            setSC3DAccNodes(lat, 0.49, Calc())
        else:
            lat = fodo_lattice(1)
            setSC2p5DrbAccNodes(lat, 0.49, Calc(), PIPE_RADIUS)
        lat.trackBunch(Stand())
    finally:
        tt.TPB = saved
    out, cur = [], []
    for c in calls:
        if c[0] == "SC":
            out.append((cur, c[1:]))
            cur = []
        else:
            cur.append((c[0],) + tuple(c[1:1 + _ORACLE_NARGS.get(c[0], len(c) - 1)]))
    return out


def cpu_baseline(args, n, grid, coords):
    """The CPU oracle on a bounded sample of the same workload: the lattice's SC intervals
    in order (every TEAPOT map of the interval, then the SC step), after one untimed
    interval that builds the Green function, until about 10 s of CPU work.  Threads: all
    host cores (OpenMP particle loops); the 3D transforms run through scipy.fft
    (pocketfft, the same threads) in place of the oracle's radix-2 FFT."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    O.set_threads(threads)
    prog = lattice_program(args.workload)
    ob = O.Bunch(coords.copy(), macrosize=1e13 / n, kin_energy=1.0)
    if args.workload == "ring3d":
        calc = O.SC3D(*grid)
        fft = O.use_library_fft3(threads)
        desc = f"SpaceChargeCalc3D{tuple(grid)}"
    else:
        calc = O.SC2p5Drb(*grid)
        fft = None
        desc = f"SpaceChargeCalc2p5Drb{tuple(grid)}"

    def interval(k):
        maps, sc = prog[k % len(prog)]
        for m in maps:
            ob.apply(*m)
        calc.trackBunch(ob, *sc)
        return len(maps)

    interval(0)
    if fft is not None:
        fft.reset()
    t0 = time.perf_counter()
    reps, nmaps = 0, 0
    while reps < 20000 and (reps < 2 or time.perf_counter() - t0 < 10.0):
        nmaps += interval(1 + reps)
        reps += 1
    total = time.perf_counter() - t0
    out = {"value": n * reps / total, "unit": "particle-pushes/s", "cores": threads,
           "kind": "port",
           "sample": (f"{n} particles, {reps} consecutive SC intervals of the lattice "
                      f"({nmaps} TEAPOT maps and {reps} {desc} steps) after one untimed "
                      f"interval"),
           "seconds": round(total, 3)}
    if fft is not None:
        out["fft"] = f"scipy.fft (pocketfft), {threads} workers"
        out["fft_share"] = round(fft.seconds / total, 3)
        O.use_library_fft3(0)
    return out


def main_batch(args):
    """64 bunches in one batch; a FODO cell captured once and replayed per step."""
    import _obnative as N
    import spacecharge
    from batch import BunchBatch
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from workloads import PIPE_RADIUS, batch_coords, batch_specs, fodo_lattice, make_bunch
    if args.workload == "fodo2p5d_graph":  # the C1 bunch alone through the graph path
        from workloads import INTENSITY, gaussian_coords
        n1 = args.n or 100_000
        specs = [{"n": n1, "grid": (128, 128, 64), "intensity": INTENSITY}]
        bunches = [make_bunch(gaussian_coords(n1), intensity=INTENSITY)]
    else:
        specs = batch_specs(64)
        bunches = [make_bunch(batch_coords(sp), intensity=sp["intensity"]) for sp in specs]
    calcs = [spacecharge.SpaceChargeCalc2p5Drb(*sp["grid"]) for sp in specs]
    bb = BunchBatch(bunches, calcs)
    lat = fodo_lattice(1)
    scn = setSC2p5DrbAccNodes(lat, 0.49, bb.calculator(), PIPE_RADIUS)
    nsc = len(scn)
    n = sum(sp["n"] for sp in specs)
    for _ in range(max(1, args.warmup)):
        lat.trackBunch(bb.bunch())  # direct cells: tuning, records
    bb.synchronize()
    bb.capture(lambda: lat.trackBunch(bb.bunch()))
    bb.replay(1)
    bb.synchronize()
    t0 = time.perf_counter()
    bb.replay(args.steps)
    bb.synchronize()
    dt = time.perf_counter() - t0
    value = n * nsc * args.steps / dt
    # kernel times from one cell tracked directly (graph replays are not instrumented)
    import ctypes as C
    N.call("ob_profile_enable", 1)
    lat.trackBunch(bb.bunch())
    bb.synchronize()
    buf = C.create_string_buffer(1 << 16)
    N.call("ob_profile_report", buf, len(buf))
    N.call("ob_profile_enable", 0)
    prof = {}
    for line in buf.value.decode().strip().splitlines():
        name, cnt, tot, mn, mx = line.split()
        prof[name] = (int(cnt), float(tot), float(mn), float(mx))
    grid_nodes = sum(sp["grid"][0] * sp["grid"][1] for sp in specs)
    kb = {"push_batch_extrema": 96 * n, "push_batch": 96 * n,
          "sc2d_hist_batch": 24 * n, "sc2d_scatter_batch": 32 * n,
          "sc2d_tdeposit_batch": 16 * n + 8 * grid_nodes}
    dom = max((k for k in prof if k in kb), key=lambda k: prof[k][1])
    cnt, tot, _, _ = prof[dom]
    avg_s = tot / cnt / 1e3
    achieved = kb[dom] / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom, "avg_us": round(avg_s * 1e6, 2), "bytes_per_launch": kb[dom]}
    kernels = {k: {"count": v[0], "avg_us": round(v[1] / v[0] * 1e3, 2),
                   "share": round(v[1] / sum(x[1] for x in prof.values()), 4)}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    cpu = None
    if not args.no_cpu and args.workload == "batch64":
        med = sorted(specs, key=lambda sp: sp["n"])[len(specs) // 2]
        a2 = argparse.Namespace(**vars(args))
        a2.workload = "fodo2p5d"
        cpu = cpu_baseline(a2, med["n"], med["grid"], batch_coords(med))
        cpu["sample"] = "median bunch of the batch: " + cpu["sample"]
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-pushes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "bunches": len(specs), "particles": n,
                   "grids": sorted({sp["grid"][0] for sp in specs}), "z_bins": 64,
                   "sc_intervals_per_step": nsc, "step": "one FODO cell (hipGraph replay)",
                   "tuning": bb.tuning(), "parallelism": "1 GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "kernels": kernels,
    }
    print(json.dumps(out))


def field_of(solver, rho, lims):
    """E = -grad phi on the nodes, phi from the device FP64 Poisson solve."""
    import numpy as np
    import spacecharge
    g = spacecharge.Grid2D(np.asarray(rho, dtype=np.float64), *lims)
    phi = spacecharge.Grid2D(rho.shape[0], rho.shape[1], *lims)
    solver.findPotential(g, phi)
    ex, ey = np.gradient(-phi.array())
    return np.stack([ex, ey])


def main_denoise(args):
    """Learned denoiser: train on (100k, 16M)-particle deposit pairs on 256 x 256 grids,
    then the field error of denoised 100k deposits against the 16M ones (FP64 solves)."""
    import numpy as np
    import _obnative as N
    import spacecharge
    from denoiser import LAYERS, Denoiser, PairSet
    H = W = 256
    n_low, n_high = args.n or 100_000, 16_000_000
    ntrain, ntest, batch = args.pairs, 64, 32
    def note(msg):
        print(f"[denoise] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    cap_pairs = capture_count(args)
    train = PairSet(H, W, capacity=ntrain + cap_pairs)
    train.generate(ntrain, n_low, n_high, seed=20261202)
    test = PairSet(H, W, ntest, n_low, n_high, seed=777)
    N.synchronize()
    t_gen = time.perf_counter() - t0
    note(f"{ntrain + ntest} generated pairs in {t_gen:.1f} s")
    t0 = time.perf_counter()
    if cap_pairs:
        capture_tracking_pairs(args, train, H, W, n_low, n_high)
    t_cap = time.perf_counter() - t0
    note(f"{train.npairs - ntrain} pairs captured from tracking in {t_cap:.1f} s")
    dn = Denoiser(H, W, max_batch=batch)
    dn.setParams(Denoiser.initParams(1))
    dn.setLoss(args.dn_scales, 1.0)
    steps_per_epoch = train.npairs // batch
    dn.train(train, epochs=1, batch=batch, lr=1e-3, npairs=batch * 4, log=None,
             shuffle=True)  # warm-up
    dn.setParams(Denoiser.initParams(1))
    N.synchronize()
    t0 = time.perf_counter()
    hist = dn.train(train, epochs=args.epochs, batch=batch, lr=1e-3, lr_final=1e-4,
                    log_every=max(1, steps_per_epoch // 4), log=note, shuffle=True, seed=9)
    N.synchronize()
    t_train = time.perf_counter() - t0
    nsteps = args.epochs * steps_per_epoch
    pix = [H * W >> (2 * lev) for _, _, lev in LAYERS]
    fwd = sum(2 * 9 * ci * co * p for (ci, co, _), p in zip(LAYERS, pix))
    dgrad = fwd - 2 * 9 * LAYERS[0][0] * LAYERS[0][1] * pix[0]
    flop_sample = fwd + dgrad + fwd
    tflops = flop_sample * batch * nsteps / t_train / 1e12
    # inference
    x, y = test.host(0, ntest)
    out = dn(x)
    xb = np.ascontiguousarray(x[:batch])
    N.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dn(xb)
    t_inf = (time.perf_counter() - t0) / 5 / batch
    # field error against the high-N deposit, FP64 device Poisson solves on a unit box
    lims = (-1.0, 1.0, -1.0, 1.0)
    solver = spacecharge.PoissonSolverFFT2D(H, W, *lims)
    ef, ep, rf, rp = [], [], [], []
    for k in range(ntest):
        ref = field_of(solver, y[k], lims)
        ef.append(np.linalg.norm(field_of(solver, out[k], lims) - ref) / np.linalg.norm(ref))
        ep.append(np.linalg.norm(field_of(solver, x[k], lims) - ref) / np.linalg.norm(ref))
        rf.append(np.linalg.norm(out[k] - y[k]) / np.linalg.norm(y[k]))
        rp.append(np.linalg.norm(x[k] - y[k]) / np.linalg.norm(y[k]))
    # cost: one 2.5D SC step at 100k particles plus inference against one at 16M
    from workloads import gaussian_coords, make_bunch
    costs = {}
    for n in (n_low, n_high):
        b = make_bunch(gaussian_coords(n, seed=5))
        calc = spacecharge.SpaceChargeCalc2p5Drb(H, W, 64)
        if n == n_low:
            calc.setDenoiser(dn)
        for _ in range(2):
            calc.trackBunch(b, 0.5, 0.1)
        N.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            calc.trackBunch(b, 0.5, 0.1)
        N.synchronize()
        costs[n] = (time.perf_counter() - t0) / 10
    tracking = track_with_denoiser(args, dn, H, W, n_low, n_high) if args.turns > 0 else None
    out_line = {
        "metric": "denoised 100k-particle field rel-L2 vs the 16M-particle FP64 solve (256x256)",
        "value": round(float(np.median(ef)), 5), "unit": "rel-L2", "n_gpus": 1,
        "steps": nsteps, "warmup": 4, "higher_is_better": False, "dtype": "bf16",
        "data": "synthetic", "vs_baseline": None,
        "config": {"workload": "denoise", "grid": [H, W], "n_low": n_low, "n_high": n_high,
                   "train_pairs": ntrain, "tracking_pairs": train.npairs - ntrain,
                   "test_pairs": ntest, "epochs": args.epochs,
                   "batch": batch, "params": dn.nparams, "loss_scales": args.dn_scales},
        "field_rel_l2": {"plain_low_n_median": round(float(np.median(ep)), 5),
                         "denoised_median": round(float(np.median(ef)), 5),
                         "plain_low_n_p90": round(float(np.percentile(ep, 90)), 5),
                         "denoised_p90": round(float(np.percentile(ef, 90)), 5)},
        "rho_rel_l2": {"plain_low_n_median": round(float(np.median(rp)), 4),
                       "denoised_median": round(float(np.median(rf)), 4)},
        "train": {"pair_generation_s": round(t_gen, 2), "pair_capture_s": round(t_cap, 2),
                  "train_s": round(t_train, 2),
                  "samples_per_s": round(batch * nsteps / t_train, 1),
                  "loss_first": hist[0][1] if hist else None,
                  "loss_last": hist[-1][1] if hist else None},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 5), "traffic": None,
                     "kernel": "training step (conv fwd + data grad + weight grad)",
                     "flop_per_sample": flop_sample},
        "inference_ms_per_grid": round(t_inf * 1e3, 4),
        "sc_step_ms": {"100k_plus_denoiser": round(costs[n_low] * 1e3, 3),
                       "16M_plain": round(costs[n_high] * 1e3, 3)},
        "tracking": tracking,
        "cpu_baseline": None,
    }
    print(json.dumps(out_line))


def capture_count(args):
    """Training pairs taken from tracking: one per SC node of args.capture_turns turns."""
    return 320 * args.capture_turns


def capture_tracking_pairs(args, pairs, H, W, n_low, n_high):
    """Track an n_high-particle Gaussian bunch (intensity 1e14, another seed than the
    tracking comparison) through the ring with the plain calculator and add a pair at
    every SC node: every (n_high / n_low)-th particle against all of them."""
    import spacecharge
    from denoiser import CapturingCalc
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from workloads import PIPE_RADIUS, gaussian_coords, make_bunch, ring_lattice
    b = make_bunch(gaussian_coords(n_high, seed=41), intensity=1e14)
    lat = ring_lattice(24)
    calc = CapturingCalc(spacecharge.SpaceChargeCalc2p5Drb(H, W, 64), pairs,
                         stride=max(1, n_high // n_low))
    setSC2p5DrbAccNodes(lat, 0.49, calc, PIPE_RADIUS)
    for _ in range(args.capture_turns):
        lat.trackBunch(b)


def track_with_denoiser(args, dn, H, W, n_low, n_high):
    """The same Gaussian bunch (intensity 1e14) through `turns` turns of the 24-cell ring
    with a SpaceChargeCalc2p5Drb(H, W, 64) node every ~0.5 m: plain at n_high (the
    reference), plain at n_low, denoised at n_low.  rms emittance growth and the fraction
    of particles beyond 4 rms in normalized transverse amplitude."""
    import numpy as np
    import _obnative as N
    import spacecharge
    from bunch import BunchTwissAnalysis
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from workloads import PIPE_RADIUS, gaussian_coords, make_bunch, ring_lattice
    res = {}
    for name, n, use in (("plain_high_n", n_high, False), ("plain_low_n", n_low, False),
                         ("denoised_low_n", n_low, True)):
        b = make_bunch(gaussian_coords(n, seed=31), intensity=1e14)
        lat = ring_lattice(24)
        calc = spacecharge.SpaceChargeCalc2p5Drb(H, W, 64)
        if use:
            calc.setDenoiser(dn)
        setSC2p5DrbAccNodes(lat, 0.49, calc, PIPE_RADIUS)
        tw = BunchTwissAnalysis()
        tw.analyzeBunch(b)
        e0 = (tw.getEmittance(0), tw.getEmittance(1))
        N.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.turns):
            lat.trackBunch(b)
        N.synchronize()
        dt = time.perf_counter() - t0
        tw.analyzeBunch(b)
        e1 = (tw.getEmittance(0), tw.getEmittance(1))
        c = b.getCoordinates()
        ax = (c[0] - c[0].mean()) / c[0].std()
        ay = (c[2] - c[2].mean()) / c[2].std()
        res[name] = {"particles": n, "eps_growth_x": e1[0] / e0[0] - 1.0,
                     "eps_growth_y": e1[1] / e0[1] - 1.0,
                     "halo_beyond_4sigma": float(np.mean(ax * ax + ay * ay > 16.0)),
                     "seconds": round(dt, 3)}
    ref = res["plain_high_n"]
    for name in ("plain_low_n", "denoised_low_n"):
        r = res[name]
        r["eps_growth_err_x"] = r["eps_growth_x"] - ref["eps_growth_x"]
        r["eps_growth_err_y"] = r["eps_growth_y"] - ref["eps_growth_y"]
    res["turns"] = args.turns
    return res


def emittances(b):
    from bunch import BunchTwissAnalysis
    tw = BunchTwissAnalysis()
    tw.analyzeBunch(b)
    return tw.getEmittance(0), tw.getEmittance(1)


def accuracy_vs_high_n(args, n, grid, ncells, e0, e1, halo):
    """The accuracy half of the metric: the same workload at 4N particles (a waterbag of the
    same seed and parameters, FP64, the same number of ring cells) is the reference.
      - rms emittance growth over the tracked cells at N and at 4N, and their difference;
      - field at one SC node: the 4N bunch after its cells and every 4th of its particles
        (an N-particle bunch of the same total charge) each get one SpaceChargeCalc3D kick;
        the kicks of the common particles (proportional to the field there) are compared,
        rel-L2 per component;
      - the fraction beyond 4 rms sizes at N and at 4N."""
    import numpy as np
    import spacecharge
    from orbit.space_charge.sc3d import setSC3DAccNodes
    from workloads import make_bunch, ring_lattice, waterbag_coords
    nh = 4 * n
    bh = make_bunch(waterbag_coords(nh))
    lat = ring_lattice(24)
    setSC3DAccNodes(lat, 0.49, spacecharge.SpaceChargeCalc3D(*grid))
    f0 = emittances(bh)
    for c in range(ncells):
        i0 = 6 * (c % 24)
        lat.trackBunch(bh, index_start=i0, index_stop=i0 + 5)
    f1 = emittances(bh)
    halo_h = beam_stats(bh)
    ch = bh.getCoordinates()
    del bh
    kicks = []
    for cc in (ch, np.ascontiguousarray(ch[:, ::4])):
        b = make_bunch(cc)
        spacecharge.SpaceChargeCalc3D(*grid).trackBunch(b, 0.5)
        kicks.append(b.getCoordinates()[[1, 3, 5]] - cc[[1, 3, 5]])
        del b
    ref = kicks[0][:, ::4]
    field = {k: float(np.linalg.norm(kicks[1][j] - ref[j]) / np.linalg.norm(ref[j]))
             for j, k in enumerate(("x", "y", "z"))}
    g = {"N": [e1[0] / e0[0] - 1.0, e1[1] / e0[1] - 1.0],
         "4N": [f1[0] / f0[0] - 1.0, f1[1] / f0[1] - 1.0]}
    g["err"] = [g["N"][0] - g["4N"][0], g["N"][1] - g["4N"][1]]
    return {"reference": f"same workload, {nh} particles, FP64", "cells": ncells,
            "eps_growth_xy": g, "field_rel_l2_vs_4N": field,
            "halo_4sigma_xy": {"N": [halo["halo_4sigma_x"], halo["halo_4sigma_y"]],
                               "4N": [halo_h["halo_4sigma_x"], halo_h["halo_4sigma_y"]]}}


def main_batch_denoise(args):
    """Batched + denoised tracking: `bunches` bunches of n particles (Gaussian, KV and
    waterbag in turn, 1e14 protons each) in one BunchBatch, a SpaceChargeCalc2p5Drb(256, 256,
    64) per bunch with one shared denoiser in every SC node, through the 24-cell ring; one
    ring cell captured in a hipGraph and replayed, one bench step per replay.  The denoiser
    is trained here first on generated (n, 16M)-particle deposit pairs of the four
    distribution families.  Accuracy: for the first `ref_bunches` bunches the same
    distributions tracked alone, plain (no denoiser) at n and at 16M particles, over the
    same cells: rms emittance growth and the fraction beyond 4 rms in normalized
    transverse amplitude."""
    import numpy as np
    import _obnative as N
    import spacecharge
    from batch import BunchBatch
    from bunch import BunchTwissAnalysis
    from denoiser import Denoiser, PairSet
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from workloads import (PIPE_RADIUS, gaussian_coords, kv_coords, make_bunch, ring_lattice,
                           waterbag_coords)
    H = W = 256
    nb, n, n_high = args.bunches, args.n or 500_000, 16_000_000

    def note(msg):
        print(f"[batch_denoise] {msg}", file=sys.stderr, flush=True)

    # z inside the 144 m ring (the ring's bunch wrap keeps z within +-72 m)
    gens = [lambda m, sd: gaussian_coords(m, seed=sd, zhalf=60.0),
            lambda m, sd: kv_coords(m, seed=sd, rx=4e-3, rxp=0.7e-3, zhalf=60.0),
            lambda m, sd: waterbag_coords(m, seed=sd, sz=20.0)]

    def coords(k, m):
        return gens[k % 3](m, 5000 + k)

    t0 = time.perf_counter()
    pairs = PairSet(H, W, capacity=args.dn_pairs)
    pairs.generate(args.dn_pairs, n, n_high, seed=20261203)
    dn = Denoiser(H, W, max_batch=32)
    dn.setParams(Denoiser.initParams(1))
    dn.setLoss(args.dn_scales, 1.0)
    hist = dn.train(pairs, epochs=args.epochs, batch=32, lr=1e-3, lr_final=1e-4, log=None,
                    shuffle=True, seed=3)
    N.synchronize()
    t_train = time.perf_counter() - t0
    note(f"{args.dn_pairs} pairs generated and trained in {t_train:.1f} s")
    del pairs

    def emit(b):
        tw = BunchTwissAnalysis()
        tw.analyzeBunch(b)
        return tw.getEmittance(0), tw.getEmittance(1)

    def halo(b):
        c = b.getCoordinates()
        ax = (c[0] - c[0].mean()) / c[0].std()
        ay = (c[2] - c[2].mean()) / c[2].std()
        return float(np.mean(ax * ax + ay * ay > 16.0))

    bunches = [make_bunch(coords(k, n), intensity=1e14) for k in range(nb)]
    calcs = [spacecharge.SpaceChargeCalc2p5Drb(H, W, 64) for _ in range(nb)]
    for c in calcs:
        c.setDenoiser(dn)
    e0 = [emit(b) for b in bunches[:args.ref_bunches]]
    bb = BunchBatch(bunches, calcs)
    lat = ring_lattice(24)
    scn = setSC2p5DrbAccNodes(lat, 0.49, bb.calculator(), PIPE_RADIUS)
    nsc = len(scn) // 24
    # every cell of the ring is the same, so one cell (nodes 0-5) captured and replayed
    # tracks the bunches through successive cells
    cell = lambda: lat.trackBunch(bb.bunch(), index_start=0, index_stop=5)
    for _ in range(max(1, args.warmup)):
        cell()
    bb.synchronize()
    graph = True
    try:
        bb.capture(cell)
        bb.replay(1)
    except Exception as e:  # the direct path is the fallback, and is reported
        note(f"graph capture failed ({e}); tracking cells directly")
        graph = False
        cell()
    bb.synchronize()
    t0 = time.perf_counter()
    if graph:
        bb.replay(args.steps)
    else:
        for _ in range(args.steps):
            cell()
    bb.synchronize()
    dt = time.perf_counter() - t0
    cells = max(1, args.warmup) + 1 + args.steps
    value = nb * n * nsc * args.steps / dt
    # kernel times of one cell tracked directly (graph replays are not instrumented)
    import ctypes as C
    N.call("ob_profile_enable", 1)
    cell()
    bb.synchronize()
    cells += 1
    buf = C.create_string_buffer(1 << 16)
    N.call("ob_profile_report", buf, len(buf))
    N.call("ob_profile_enable", 0)
    prof = {}
    for line in buf.value.decode().strip().splitlines():
        name, cnt, tot, mn, mx = line.split()
        prof[name] = (int(cnt), float(tot), float(mn), float(mx))
    kernels = {k: {"count": v[0], "avg_us": round(v[1] / v[0] * 1e3, 2),
                   "share": round(v[1] / sum(x[1] for x in prof.values()), 4)}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    ntot = nb * n
    kb = {"push_batch_extrema": 96 * ntot, "push_batch": 96 * ntot}
    dom = max((k for k in prof if k in kb), key=lambda k: prof[k][1])
    cnt, tot, _, _ = prof[dom]
    avg_s = tot / cnt / 1e3
    achieved = kb[dom] / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom, "avg_us": round(avg_s * 1e6, 2), "bytes_per_launch": kb[dom]}
    res = []
    for k in range(args.ref_bunches):
        e1 = emit(bunches[k])
        row = {"distribution": ("gaussian", "kv", "waterbag")[k % 3],
               "denoised_n": {"particles": n, "eps_growth_xy": [e1[0] / e0[k][0] - 1,
                                                                e1[1] / e0[k][1] - 1],
                              "halo_4sigma": halo(bunches[k])}}
        for name, m in (("plain_n", n), ("plain_16M", n_high)):
            b = make_bunch(coords(k, m), intensity=1e14)
            c = spacecharge.SpaceChargeCalc2p5Drb(H, W, 64)
            lr = ring_lattice(24)
            setSC2p5DrbAccNodes(lr, 0.49, c, PIPE_RADIUS)
            f0 = emit(b)
            for j in range(cells):
                i0 = 6 * (j % 24)
                lr.trackBunch(b, index_start=i0, index_stop=i0 + 5)
            f1 = emit(b)
            row[name] = {"particles": m, "eps_growth_xy": [f1[0] / f0[0] - 1,
                                                           f1[1] / f0[1] - 1],
                         "halo_4sigma": halo(b)}
            del b, c
        ref = row["plain_16M"]["eps_growth_xy"]
        for name in ("denoised_n", "plain_n"):
            g = row[name]["eps_growth_xy"]
            row[name]["eps_growth_err_xy"] = [g[0] - ref[0], g[1] - ref[1]]
        res.append(row)
        note(f"reference bunch {k} done")
    out = {
        "metric": "particle-pushes/s incl. SC kick (batched, denoised), 1 MI355X",
        "value": round(value, 1), "unit": "particle-pushes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (denoiser bf16)",
        "data": "synthetic",
        "config": {"workload": "batch_denoise", "bunches": nb, "particles_per_bunch": n,
                   "grid": [H, W, 64], "lattice": "24-cell ring (144 m), one cell per step",
                   "sc_intervals_per_step": nsc, "graph": graph,
                   "tuning": bb.tuning(),
                   "train_pairs": args.dn_pairs, "epochs": args.epochs,
                   "intensity": 1e14,
                   "train_s": round(t_train, 1), "parallelism": "1 GPU"},
        "roofline": roofline, "kernels": kernels,
        "accuracy": {"cells": cells, "reference": "each bunch alone, plain FP64 at 16M",
                     "bunches": res},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    args = parse()
    if args.workload == "batch_denoise" and args.gpus == 1:
        return main_batch_denoise(args)
    if args.workload in ("batch64", "fodo2p5d_graph") and args.gpus == 1:
        return main_batch(args)
    if args.workload == "denoise" and args.gpus == 1:
        return main_denoise(args)
    if args.gpus != 1:
        print("bench.py: only one GPU is measured; runs on several GPUs are not supported "
              "(the path would shard by particles with one rho all-reduce per SC step)",
              file=sys.stderr)
        sys.exit(2)
    import _obnative as N
    b, lat, calc, cells, n, grid, coords = build_workload(args)
    e0 = emittances(b)
    N.synchronize()
    run_cells(lat, b, cells, 0, args.warmup)
    N.synchronize()
    t0 = time.perf_counter()
    pushes_sc = run_cells(lat, b, cells, args.warmup, args.steps)
    N.synchronize()
    dt = time.perf_counter() - t0
    value = n * pushes_sc / dt

    prof = profile(N, lat, b, cells, args.profile_steps)
    kb = kernel_bytes(args.workload, n, grid)
    dom = max((k for k in prof if k in kb), key=lambda k: prof[k][1])
    cnt, tot, _, _ = prof[dom]
    avg_s = tot / cnt / 1e3
    achieved = kb[dom] / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(args.workload, dom),
                "kernel": dom, "avg_us": round(avg_s * 1e6, 2), "bytes_per_launch": kb[dom]}
    kernels = {k: {"count": v[0], "avg_us": round(v[1] / v[0] * 1e3, 2),
                   "share": round(v[1] / sum(x[1] for x in prof.values()), 4)}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    e1 = emittances(b)
    beam = beam_stats(b)
    acc = None
    if args.workload == "ring3d" and not args.no_accuracy:
        acc = accuracy_vs_high_n(args, n, grid, args.warmup + args.steps + args.profile_steps,
                                 e0, e1, beam)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, n, grid, coords)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": n, "grid": list(grid),
                   "sc_intervals_timed": pushes_sc,
                   "step": "one ring cell" if args.workload == "ring3d" else "one FODO cell",
                   "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "beam": dict(beam, emit0_x=e0[0], emit0_y=e0[1]),
        "accuracy": acc,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
