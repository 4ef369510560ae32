"""Per-kernel resource report (VGPR, AGPR, SGPR, scratch, LDS, occupancy) for gfx950.

AMD counterpart of the reference's ptxas --resource-usage table
(Tridiagonal_CUDA/resource_usage_format.cpp:32-94, Makefile:24-29): it compiles a
source with -Rpass-analysis=kernel-resource-usage and tabulates the remarks.

    python -m tridiagonalsolver_amd.resource_usage [csrc/file.hip ...] [-o out.txt]
"""
import os
import re
import subprocess
import sys

from . import build as _b

# remark labels (older compilers print "SGPRs", newer ones "TotalSGPRs")
_FIELDS = [("VGPRs", "vgpr"), ("AGPRs", "agpr"), ("SGPRs", "sgpr"), ("TotalSGPRs", "sgpr"),
           ("VGPRs Spill", "vspill"), ("SGPRs Spill", "sspill"),
           ("ScratchSize [bytes/lane]", "scratch"), ("Occupancy [waves/SIMD]", "occ"),
           ("LDS Size [bytes/block]", "lds")]


def _demangle(names):
    try:
        r = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True)
        return r.stdout.splitlines()
    except OSError:
        return names


def report(src):
    cmd = [_b.HIPCC] + _b.HIPFLAGS + ["-c", src, "-o", os.devnull,
                                      "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    rows, cur = [], None
    for line in r.stderr.splitlines():
        m = re.search(r"remark:\s+Function Name:\s+(\S+)", line)
        if m:
            cur = {"name": m.group(1)}
            rows.append(cur)
            continue
        if cur is None:
            continue
        for label, key in _FIELDS:
            m = re.search(r"remark:\s+%s:\s+(\d+)" % re.escape(label), line)
            if m:
                cur[key] = int(m.group(1))
    names = _demangle([row["name"] for row in rows])
    for row, nm in zip(rows, names):
        row["name"] = nm
    return rows


def format_rows(rows):
    out = ["%-6s %-6s %-6s %-7s %-7s %-8s %-4s %-7s  %s" % (
        "VGPR", "AGPR", "SGPR", "vspill", "sspill", "scratch", "occ", "LDS", "kernel")]
    for r in rows:
        out.append("%-6s %-6s %-6s %-7s %-7s %-8s %-4s %-7s  %s" % (
            r.get("vgpr", "?"), r.get("agpr", "?"), r.get("sgpr", "?"), r.get("vspill", "?"),
            r.get("sspill", "?"), r.get("scratch", "?"), r.get("occ", "?"), r.get("lds", "?"),
            r["name"]))
    return "\n".join(out)


def main(argv):
    out = None
    if "-o" in argv:
        i = argv.index("-o")
        out = argv[i + 1]
        argv = argv[:i] + argv[i + 2:]
    srcs = argv or sorted(os.path.join(_b.CSRC, f) for f in os.listdir(_b.CSRC)
                          if f.endswith(".hip"))
    text = "\n\n".join("== %s\n%s" % (os.path.basename(s), format_rows(report(s))) for s in srcs)
    if out:
        with open(out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main(sys.argv[1:])
