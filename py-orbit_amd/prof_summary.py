"""Summaries of rocprofv3 result databases (rocpd SQLite) for this library's kernels.

    python py-orbit_amd/prof_summary.py stats  <results.db>
    python py-orbit_amd/prof_summary.py pmc    <results.db> <copy_results.db> <mib>
    python py-orbit_amd/prof_summary.py traffic <fetch.db> <copy_fetch.db> <write.db> \
                                                <copy_write.db> <mib>

`stats` prints per-kernel calls, total/avg/min/max duration and share (the content of
`rocprofv3 --stats`).  `pmc` prints the mean per-dispatch value of each counter per kernel,
and the same value scaled by the copy yardstick: copy_kernel (py-orbit_amd/
calibrate_copy.py) moves exactly <mib> MiB each way per dispatch, so
scale = <mib>*1024 KiB / (its counter reading).  `traffic` joins a FETCH_SIZE and a
WRITE_SIZE run into scaled HBM bytes per dispatch, one `traffic_bytes <kernel> <fetch>
<write>` line per kernel (the form bench.py reads for its roofline entry).
"""
import sqlite3
import sys


def short(name):
    n = name.replace("(anonymous namespace)::", "")
    n = n.split("(")[0]
    for pre in ("void ", "ob::"):
        n = n.replace(pre, "")
    return n.strip().replace(", ", ",")


def kernel_rows(db):
    c = sqlite3.connect(db).cursor()
    rows = c.execute("select name, duration from kernels").fetchall()
    agg = {}
    for name, dur in rows:
        k = short(name)
        a = agg.setdefault(k, [0, 0.0, 1e30, 0.0])
        a[0] += 1
        a[1] += dur
        a[2] = min(a[2], dur)
        a[3] = max(a[3], dur)
    return agg


def stats(db):
    agg = kernel_rows(db)
    tot = sum(v[1] for v in agg.values())
    print(f"{'kernel':40s} {'calls':>6s} {'total_ms':>10s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'share':>7s}")
    for k, (n, t, mn, mx) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"{k:40s} {n:6d} {t / 1e6:10.3f} {t / n / 1e3:10.2f} {mn / 1e3:10.2f} {mx / 1e3:10.2f} {t / tot:7.2%}")


def pmc_means(db):
    c = sqlite3.connect(db).cursor()
    rows = c.execute("select name, counter_name, counter_value from pmc_events").fetchall()
    agg = {}
    for name, cn, v in rows:
        a = agg.setdefault((short(name), cn), [0, 0.0])
        a[0] += 1
        a[1] += v
    return {k: v[1] / v[0] for k, v in agg.items()}


def pmc(db, copy_db, mib):
    m = pmc_means(db)
    cm = pmc_means(copy_db)
    scale = {}
    for (k, cn), v in cm.items():
        if k == "copy_kernel":
            scale[cn] = mib * 1024.0 / v
    print(f"copy_kernel yardstick ({mib} MiB each way): " +
          ", ".join(f"{cn} reads {mib * 1024 / s:.0f} KiB -> scale {s:.4f}" for cn, s in scale.items()))
    print(f"{'kernel':40s} {'counter':>12s} {'KiB/dispatch':>14s} {'scaled_MB':>12s}")
    for (k, cn), v in sorted(m.items()):
        s = scale.get(cn, 1.0)
        print(f"{k:40s} {cn:>12s} {v:14.1f} {v * s * 1024 / 1e6:12.2f}")


def scaled(db, copy_db, mib, counter):
    m = pmc_means(db)
    cm = pmc_means(copy_db)
    s = mib * 1024.0 / cm[("copy_kernel", counter)]
    return s, {k: v * s * 1024 for (k, cn), v in m.items() if cn == counter}


def traffic(fdb, cfdb, wdb, cwdb, mib):
    sf, f = scaled(fdb, cfdb, mib, "FETCH_SIZE")
    sw, w = scaled(wdb, cwdb, mib, "WRITE_SIZE")
    print(f"scale: FETCH_SIZE x{sf:.4f}, WRITE_SIZE x{sw:.4f} (copy_kernel, {mib:.0f} MiB each way)")
    print(f"{'kernel':40s} {'fetch_MB':>10s} {'write_MB':>10s} {'total_MB':>10s}")
    for k in sorted(f, key=lambda k: -(f[k] + w.get(k, 0.0))):
        print(f"{k:40s} {f[k] / 1e6:10.2f} {w.get(k, 0.0) / 1e6:10.2f} {(f[k] + w.get(k, 0.0)) / 1e6:10.2f}")
    for k in sorted(f):
        print(f"traffic_bytes {k} {f[k]:.0f} {w.get(k, 0.0):.0f}")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif sys.argv[1] == "pmc":
        pmc(sys.argv[2], sys.argv[3], float(sys.argv[4]))
    elif sys.argv[1] == "traffic":
        traffic(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], float(sys.argv[6]))
