"""Per-kernel summary of rocprofv3 output (rocpd SQLite databases).

Usage::

    python -m anuga_cuda_amd.rocprof_summary --trace T.db [--fetch F.db --write W.db]
        [--copy-bytes B] [--triangles N] [--json out.json]

* ``--trace``: a ``rocprofv3 --kernel-trace`` database; gives calls and the
  average / min / max duration per kernel.
* ``--fetch`` / ``--write``: databases of ``rocprofv3 --pmc FETCH_SIZE`` and
  ``--pmc WRITE_SIZE`` runs (KiB per dispatch).  Both counters are scaled by
  what they read for the largest plain device copy in the same run
  (``__amd_rocclr_copyBuffer`` of ``--copy-bytes`` bytes read and written):
  scale = copy bytes / counter bytes.  On gfx950 FETCH_SIZE counts the 128-byte
  read requests of the memory fabric at a smaller size, so its factor is not 1.
* ``--json``: writes {kernel short name: {avg_us, fetch_bytes, write_bytes,
  traffic_bytes}} for bench.py's roofline ``traffic`` field.
* ``--pmc DB`` (repeatable): databases of other ``rocprofv3 --pmc`` runs;
  prints per kernel, averaged over its dispatches, what the SQ / TCC counters
  present in them give: VALU instructions, FP64 transcendental instructions
  and vector memory reads per wavefront (SQ_INSTS_* / SQ_WAVES), the share
  of wavefront cycles spent waiting (SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES) and
  issuing VALU work (SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES), and the L2 hit
  rate (TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum)).
* ``--bytes NAME=B`` (repeatable): algorithmic bytes per launch of the
  kernel whose base name (the function name without namespaces, template or
  parameter lists; mangled names are parsed) is NAME; prints its achieved
  GB/s (B / average duration) against the 8 TB/s HBM peak.  ``A+B=bytes``
  names an operation of two kernels launched once each (their average
  durations add).
"""
from __future__ import annotations

import argparse
import json
import sqlite3
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name):
    n = name.replace("(anonymous namespace)::", "").split("(")[0]
    if n.startswith("void "):
        n = n[5:]
    return n.replace("sw::", "")


def base_name(name):
    """Function name of a kernel symbol without namespaces, template
    arguments or parameters: '_ZN3gnn11k_mp_directE...' -> 'k_mp_direct',
    'gnn::k_apply' -> 'k_apply', 'void k_flux<double, true, 256>(...)' ->
    'k_flux'."""
    if name.startswith("_ZN"):
        i, parts = 3, []
        while i < len(name) and name[i].isdigit():
            j = i
            while j < len(name) and name[j].isdigit():
                j += 1
            n = int(name[i:j])
            parts.append(name[j:j + n])
            i = j + n
        if parts:
            return parts[-1]
    n = short_name(name).split("<")[0]
    return n.split("::")[-1].strip()


def kernel_stats(db):
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), avg(duration), min(duration), max(duration) from kernels group by name").fetchall()
    return {r[0]: (r[1], r[2] / 1e3, r[3] / 1e3, r[4] / 1e3) for r in rows}


def counter(db, name):
    """{kernel: (dispatches, mean value, max value)} of one PMC counter."""
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), avg(counter_value), max(counter_value) from pmc_events "
        "where counter_name = ? group by name", (name,)).fetchall()
    return {r[0]: (r[1], r[2], r[3]) for r in rows}


def counters_by_dispatch(db):
    """{kernel: {counter: mean over dispatches of the counter's sum over
    its instances (shader engines, channels)}}."""
    con = sqlite3.connect(db)
    per = {}
    for name, cname, disp, val in con.execute(
            "select name, counter_name, dispatch_id, sum(counter_value) from pmc_events "
            "group by name, counter_name, dispatch_id"):
        per.setdefault(name, {}).setdefault(cname, []).append(val)
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in per.items()}


def pmc_figures(c):
    """Derived figures of one kernel's counters (those present)."""
    out = {}
    w = c.get("SQ_WAVES")
    for key, label in (("SQ_INSTS_VALU", "valu_per_wave"), ("SQ_INSTS_VALU_TRANS_F64", "trans_f64_per_wave"),
                       ("SQ_INSTS_VMEM_RD", "vmem_rd_per_wave"), ("SQ_INSTS_MFMA", "mfma_per_wave"),
                       ("SQ_INSTS_LDS", "lds_per_wave")):
        if w and key in c:
            out[label] = c[key] / w
    cyc = c.get("SQ_WAVE_CYCLES")
    if cyc:
        if "SQ_WAIT_INST_ANY" in c:
            out["wait_frac"] = c["SQ_WAIT_INST_ANY"] / cyc
        if "SQ_ACTIVE_INST_VALU" in c:
            out["valu_active_frac"] = c["SQ_ACTIVE_INST_VALU"] / cyc
        if "SQ_WAIT_ANY" in c:
            out["wait_any_frac"] = c["SQ_WAIT_ANY"] / cyc
    if "SQ_LDS_BANK_CONFLICT" in c and c.get("SQ_ACTIVE_INST_LDS"):
        out["lds_conflict_frac"] = c["SQ_LDS_BANK_CONFLICT"] / c["SQ_ACTIVE_INST_LDS"]
    if "TCC_HIT_sum" in c and "TCC_MISS_sum" in c:
        tot = c["TCC_HIT_sum"] + c["TCC_MISS_sum"]
        if tot:
            out["l2_hit"] = c["TCC_HIT_sum"] / tot
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--triangles", type=int, default=0, help="triangles per launch, for bytes per triangle")
    ap.add_argument("--json")
    ap.add_argument("--bytes", action="append", default=[], metavar="NAME=B")
    ap.add_argument("--pmc", action="append", default=[], metavar="DB")
    a = ap.parse_args(argv)

    ks = kernel_stats(a.trace)
    out = {}
    print("%-44s %6s %10s %10s %10s" % ("kernel", "calls", "avg_us", "min_us", "max_us"))
    for name, (n, avg, mn, mx) in sorted(ks.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        print("%-44s %6d %10.2f %10.2f %10.2f" % (short_name(name)[:44], n, avg, mn, mx))
        out[short_name(name)] = {"calls": n, "avg_us": avg}
    if a.fetch and a.write:
        f = counter(a.fetch, "FETCH_SIZE")
        w = counter(a.write, "WRITE_SIZE")
        if COPY_KERNEL not in f or COPY_KERNEL not in w:
            sys.exit("no %s dispatch in the PMC runs: run the workload with a device copy" % COPY_KERNEL)
        # the largest copy dispatch is the calibration copy
        sf = a.copy_bytes / (f[COPY_KERNEL][2] * 1024.0)
        sw = a.copy_bytes / (w[COPY_KERNEL][2] * 1024.0)
        print()
        print("calibration: copy of %.0f MiB reads FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB" % (
            a.copy_bytes / 2**20, f[COPY_KERNEL][2], w[COPY_KERNEL][2]))
        print("scale factors: FETCH_SIZE x %.4f, WRITE_SIZE x %.4f" % (sf, sw))
        print()
        hdr = "%-44s %10s %10s %10s" % ("kernel", "fetch_MB", "write_MB", "total_MB")
        if a.triangles:
            hdr += " %10s" % "B/triangle"
        print(hdr)
        for name in sorted(f):
            if name not in w or name == COPY_KERNEL:
                continue
            fb = f[name][1] * 1024.0 * sf
            wb = w[name][1] * 1024.0 * sw
            line = "%-44s %10.2f %10.2f %10.2f" % (short_name(name)[:44], fb / 1e6, wb / 1e6, (fb + wb) / 1e6)
            if a.triangles:
                line += " %10.1f" % ((fb + wb) / a.triangles)
            print(line)
            d = out.setdefault(short_name(name), {})
            d.update({"fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb})
    if a.bytes:
        print()
        print("%-44s %12s %10s %10s %8s" % ("kernel", "alg_bytes", "avg_us", "GB/s", "of_8TB/s"))
        for spec in a.bytes:
            names, b = spec.split("=")
            parts = names.split("+")
            found = {p: [(nm, v) for nm, v in ks.items() if base_name(nm) == p] for p in parts}
            if not all(found.values()):
                print("%-44s (no kernel named %s)" % (names[:44], [p for p in parts if not found[p]]))
                continue
            if len(parts) == 1:
                rows = [(short_name(nm), v[1]) for nm, v in sorted(found[parts[0]])]
            else:
                # one launch of each kernel per operation: average durations add
                rows = [(names, sum(sum(v[1] for _, v in found[p]) / len(found[p]) for p in parts))]
            for label, avg in rows:
                gbs = float(b) / (avg * 1e-6) / 1e9
                print("%-44s %12.4g %10.2f %10.1f %8.3f" % (label[:44], float(b), avg, gbs, gbs / 8000.0))
                out.setdefault(label, {})["algorithmic_bytes"] = float(b)
    if a.pmc:
        merged = {}
        for db in a.pmc:
            for name, c in counters_by_dispatch(db).items():
                merged.setdefault(short_name(name), {}).update(c)
        figs = {name: pmc_figures(c) for name, c in merged.items()}
        cols = [c for c in ("valu_per_wave", "trans_f64_per_wave", "vmem_rd_per_wave", "mfma_per_wave",
                            "lds_per_wave", "wait_frac", "wait_any_frac", "valu_active_frac", "lds_conflict_frac",
                            "l2_hit") if any(c in f for f in figs.values())]
        print()
        print("%-44s" % "kernel" + "".join(" %12s" % c[:12] for c in cols))
        for name in sorted(figs):
            if name.startswith("__amd") or not figs[name]:
                continue
            print("%-44s" % name[:44] + "".join(
                " %12.3f" % figs[name][c] if c in figs[name] else " %12s" % "-" for c in cols))
            out.setdefault(name, {})["counters"] = figs[name]
    if a.json:
        with open(a.json, "w") as fh:
            json.dump({"fetch_scale": sf if a.fetch else None, "write_scale": sw if a.write else None,
                       "kernels": out}, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
