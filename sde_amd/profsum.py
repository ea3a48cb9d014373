"""Summary of rocprofv3 runs of bench.py: per-kernel duration and HBM traffic.

    python -m sde_amd.profsum OUT "<command line>" > profiles/<name>.md

OUT holds three rocprofv3 output folders of the same command, and others
for the LDS and VALU counters when present:
    OUT/trace  --kernel-trace --stats
    OUT/fetch  --pmc FETCH_SIZE
    OUT/write  --pmc WRITE_SIZE
    OUT/lds    --pmc SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE SQ_LDS_ADDR_CONFLICT
    OUT/valu   --pmc VALUBusy SQ_INSTS_VALU SQ_INSTS_SALU GRBM_GUI_ACTIVE
each written with ``-o run`` (rocpd SQLite databases, run_results.db).

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are calibrated against the
plain device copy that ``bench.py --copy-probe`` launches (k_copy, 1 GiB
read + 1 GiB written): factor = 1 GiB / counter reading of that copy.  The
table gives, per kernel, the dispatch count, the average duration, the
calibrated read and written megabytes per dispatch and the resulting
bandwidth, and that bandwidth as a fraction of MI355X's 8 TB/s HBM peak.
bench.py reads the "hbm_bytes" column for its roofline.  The counters
count the traffic between L2 and the fabric, so bytes served by the
Infinity Cache (MALL) are included: a kernel can show more than the copy's
rate.

The LDS table gives, per kernel, the LDS instructions per dispatch, the
cycles lost to bank conflicts per LDS instruction and the shares of the
LDS-active cycles (SQ_LDS_IDX_ACTIVE) spent in bank and in address
conflicts.

The VALU table gives, per kernel, the vector and scalar ALU instructions
per dispatch (wave instructions), VALUBusy (the share of the kernel's GPU
time the CUs spend executing VALU instructions, in percent) and the VALU
instructions issued per CU per cycle of the kernel's GPU time
(SQ_INSTS_VALU / (CUs * GRBM_GUI_ACTIVE per XCD)); one wave64 FP32
instruction per CU per cycle is the issue rate of the four SIMDs.
"""
import collections
import glob
import os
import re
import sqlite3
import sys


def short_name(name):
    name = re.sub(r"\(anonymous namespace\)::", "", name)
    name = re.sub(r"^void ", "", name)
    if name.startswith("rocprim::"):  # rocprim::<ns>::detail::trampoline_kernel<...>
        return "rocprim::" + re.sub(r"<.*", "", name).split("::")[-1]
    return name.split("(")[0].strip()


def labels(names):
    """Row label of every full kernel name: its short name, or, where several
    kernels share one short name (rocPRIM's histogram and onesweep kernels
    are all one trampoline template), the short name with "[k of n]" in the
    order of the sorted full names, the same in every database of a run."""
    groups = collections.defaultdict(list)
    for n in sorted(set(names)):
        groups[short_name(n)].append(n)
    out = {}
    for s, full in groups.items():
        for k, n in enumerate(full):
            out[n] = s if len(full) == 1 else "%s [%d of %d]" % (s, k + 1, len(full))
    return out


def _db(folder):
    hits = glob.glob(os.path.join(folder, "**", "*_results.db"), recursive=True)
    if not hits:
        raise SystemExit("no rocpd database under %s" % folder)
    return hits[0]


def trace(folder):
    c = sqlite3.connect(_db(folder))
    rows = c.execute("select name, count(*), avg(end - start) from kernels group by name").fetchall()
    lab = labels([r[0] for r in rows])
    return {lab[r[0]]: (r[1], r[2] / 1e6) for r in rows}


def counters(folder):
    c = sqlite3.connect(_db(folder))
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    rows = c.execute("select kernel_name, counter_name, value from counters_collection").fetchall()
    lab = labels([r[0] for r in rows])
    for name, ctr, val in rows:
        acc[lab[name]][ctr].append(val)
    return {k: {ctr: sum(v) / len(v) for ctr, v in d.items()} for k, d in acc.items()}


CUS = 256  # MI355X: 8 XCDs of 32 CUs
XCDS = 8


def summarize(out, command):
    t = trace(os.path.join(out, "trace"))
    f = counters(os.path.join(out, "fetch"))
    w = counters(os.path.join(out, "write"))
    cf = f.get("k_copy", {}).get("FETCH_SIZE")
    cw = w.get("k_copy", {}).get("WRITE_SIZE")
    kf = 1048576.0 / cf if cf else 1.0
    kw = 1048576.0 / cw if cw else 1.0
    lines = ["# Kernel durations and HBM traffic", "", "Command (from the repository root):", "", "```"]
    lines += [command, "```", ""]
    lines.append("Copy calibration (k_copy, 1 GiB read + 1 GiB written per dispatch): FETCH_SIZE %.0f KiB, "
                 "WRITE_SIZE %.0f KiB; factors fetch x%.3f, write x%.3f." % (cf or 0, cw or 0, kf, kw))
    lines.append("")
    lines.append("| kernel | dispatches | avg_ms | read_MB | written_MB | hbm_bytes | GB/s | frac_8TBs |")
    lines.append("|---|---|---|---|---|---|---|---|")
    order = sorted(t.items(), key=lambda kv: -kv[1][0] * kv[1][1])
    for k, (n, avg) in order:
        rd = f.get(k, {}).get("FETCH_SIZE", 0.0) * kf * 1024.0
        wr = w.get(k, {}).get("WRITE_SIZE", 0.0) * kw * 1024.0
        gbs = (rd + wr) / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
        lines.append("| %s | %d | %.4f | %.1f | %.1f | %.0f | %.0f | %.3f |"
                     % (k, n, avg, rd / 1e6, wr / 1e6, rd + wr, gbs, gbs / 8000.0))
    lds_dir = os.path.join(out, "lds")
    if os.path.isdir(lds_dir):
        l = counters(lds_dir)
        lines += ["", "LDS counters (per dispatch):", "",
                  "| kernel | lds_instr | bank_conflict_cycles | conflict_cycles_per_instr | bank_share | addr_share |",
                  "|---|---|---|---|---|---|"]
        for k, _ in order:
            c = l.get(k, {})
            ni, bc = c.get("SQ_INSTS_LDS", 0.0), c.get("SQ_LDS_BANK_CONFLICT", 0.0)
            act, ac = c.get("SQ_LDS_IDX_ACTIVE", 0.0), c.get("SQ_LDS_ADDR_CONFLICT", 0.0)
            if ni <= 0:
                continue
            lines.append("| %s | %.0f | %.0f | %.3f | %.3f | %.3f |"
                         % (k, ni, bc, bc / ni, bc / act if act else 0.0, ac / act if act else 0.0))
    valu_dir = os.path.join(out, "valu")
    if os.path.isdir(valu_dir):
        v = counters(valu_dir)
        lines += ["", "VALU counters (per dispatch):", "",
                  "| kernel | valu_instr | salu_instr | valu_busy_pct | valu_per_cu_cycle |",
                  "|---|---|---|---|---|"]
        for k, _ in order:
            c = v.get(k, {})
            nv, ns = c.get("SQ_INSTS_VALU", 0.0), c.get("SQ_INSTS_SALU", 0.0)
            busy, gui = c.get("VALUBusy", 0.0), c.get("GRBM_GUI_ACTIVE", 0.0)
            if nv <= 0:
                continue
            ipc = nv / (CUS * gui / XCDS) if gui > 0 else 0.0
            lines.append("| %s | %.0f | %.0f | %.1f | %.3f |" % (k, nv, ns, busy, ipc))
    return "\n".join(lines) + "\n"


def read_table(path):
    """Rows of a summary written by summarize(): kernel -> dict of columns."""
    out = {}
    if not os.path.exists(path):
        return out
    seen = False  # the first table only
    with open(path) as fh:
        for line in fh:
            if not line.startswith("|"):
                if seen:
                    break
                continue
            seen = True
            p = [x.strip() for x in line.strip().strip("|").split("|")]
            if len(p) < 7 or p[0] in ("kernel", "---") or p[0].startswith("---"):
                continue
            try:
                out[p[0]] = {"dispatches": int(p[1]), "avg_ms": float(p[2]), "read_MB": float(p[3]),
                             "written_MB": float(p[4]), "hbm_bytes": float(p[5])}
            except ValueError:
                continue
    return out


if __name__ == "__main__":
    if len(sys.argv) < 3:
        raise SystemExit(__doc__)
    sys.stdout.write(summarize(sys.argv[1], sys.argv[2]))
