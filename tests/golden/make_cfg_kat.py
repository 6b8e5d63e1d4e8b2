"""Writes tests/golden/cfg_kat.json: known-answer data for the v1.0 config
parser, derived from the reference's shipped per-GPU tables
(config/hml_sgemm_*.cfg, config/hml_sgemv_*.cfg) with an independent Python
parse of the grammar (src/hml_sgemm_kernel.cu:573-602,
src/hml_sgemv_kernel.cu:650-679: '#' lines skipped, "%d %d : %d %d" /
"%d : %d %d %d", later duplicates overwrite earlier ones).

Only derived data is stored (the resolved table entries and summary
counts), not the files.  Run in a checkout that has the reference mounted:
    python tests/golden/make_cfg_kat.py /root/reference/config
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

SGEMM_RE = re.compile(r"^\s*(-?\d+)\s+(-?\d+)\s*:\s*(-?\d+)\s+(-?\d+)")
SGEMV_RE = re.compile(r"^\s*(-?\d+)\s*:\s*(-?\d+)\s+(-?\d+)\s+(-?\d+)")


def parse(path, kind):
    table = {}
    lines = 0
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            m = (SGEMM_RE if kind == "sgemm" else SGEMV_RE).match(line)
            if not m:
                continue
            lines += 1
            v = [int(t) for t in m.groups()]
            key = tuple(v[:2]) if kind == "sgemm" else (v[0],)
            table[key] = v[2:] if kind == "sgemm" else v[1:]
    return table, lines


def main(cfgdir):
    out = {}
    for name in sorted(os.listdir(cfgdir)):
        if not name.endswith(".cfg") or "header" in name:
            continue
        kind = "sgemm" if name.startswith("hml_sgemm") else "sgemv"
        table, lines = parse(os.path.join(cfgdir, name), kind)
        hist = {}
        for val in table.values():
            hist[str(val[0])] = hist.get(str(val[0]), 0) + 1
        out[name] = {
            "kind": kind,
            "rule_lines": lines,
            "unique_keys": len(table),
            "type_histogram": hist,
            "entries": sorted([list(k) + list(v) for k, v in table.items()]),
        }
    with open(os.path.join(HERE, "cfg_kat.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/config")
