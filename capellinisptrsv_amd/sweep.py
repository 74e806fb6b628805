"""Sweep the drop-in program over a set of Matrix Market files.

The reference sweeps its matrix collection with one shell loop per program
(openCL/WrtFst/run.sh, openCL/mix/run.sh, ...): remove result.csv, run the
program on every file of a directory, each run appending one row.  The
matrix list of the paper's collection is matrix/matrix_name.txt (one path
per line).  This driver does the same with the gfx950 program:

    python -m capellinisptrsv_amd.sweep DIR_OR_LIST [--root DIR] [-a wrtfst,mix,2part,csc,auto,local,partition,serial]
           [--f32] [--values] [--out-dir DIR] [--keep] [--dry-run]

DIR_OR_LIST is a directory (every *.mtx in it, sorted) or a text file of
paths; a listed path that does not exist is looked up by file name under
--root.  Every (matrix, algorithm) pair runs `sptrsv_main` (or
`sptrsv_main_f32`) in its own child process started in --out-dir (default:
the working directory), which appends one row to result.csv there and one
JSON object to result.jsonl (the FP64 serial-solve error and the roofline
figures of the run); both files are removed first unless --keep.  "serial"
times the reference's serial CPU executor into the same columns.  Exit status:
0 if every run succeeded, 1 otherwise (failures are listed).
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ALGOS = ("wrtfst", "2part", "mix", "csc", "auto", "local", "partition", "level", "serial")


def matrices(source: str, root: str | None = None) -> list[str]:
    """The .mtx files of a directory (sorted) or of a list file (in order)."""
    if os.path.isdir(source):
        return [os.path.abspath(os.path.join(source, f)) for f in sorted(os.listdir(source)) if f.endswith(".mtx")]
    out = []
    with open(source) as f:
        for line in f:
            p = line.strip()
            if not p or p.startswith("#"):
                continue
            if not os.path.exists(p) and root:
                cand = os.path.join(root, os.path.basename(p))
                if os.path.exists(cand):
                    p = cand
            out.append(os.path.abspath(p))
    return out


def commands(mtx: list[str], algos: list[str], f32: bool = False, values: bool = False) -> list[list[str]]:
    exe = os.path.join(HERE, "sptrsv_main_f32" if f32 else "sptrsv_main")
    cmds = []
    for path in mtx:
        for a in algos:
            c = [exe, path, "-a", a]
            if values:
                c.append("--values")
            cmds.append(c)
    return cmds


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("source", help="directory of .mtx files or a list file (e.g. matrix_name.txt)")
    ap.add_argument("--root", help="directory to look listed files up in by name")
    ap.add_argument("-a", "--algos", default="wrtfst", help=f"comma-separated subset of {','.join(ALGOS)}")
    ap.add_argument("--f32", action="store_true", help="VALUE_TYPE float build of the program")
    ap.add_argument("--values", action="store_true", help="keep the matrix values (default: unit values)")
    ap.add_argument("--out-dir", default=".", help="where result.csv and result.jsonl are written")
    ap.add_argument("--keep", action="store_true", help="append to an existing CSV")
    ap.add_argument("--dry-run", action="store_true", help="print the commands only")
    ap.add_argument("--timeout", type=float, default=600.0, help="seconds per run")
    a = ap.parse_args(argv)
    algos = [x for x in a.algos.split(",") if x]
    bad = [x for x in algos if x not in ALGOS]
    if bad:
        ap.error(f"unknown algorithm(s) {bad}")
    mtx = matrices(a.source, a.root)
    cmds = commands(mtx, algos, a.f32, a.values)
    if a.dry_run:
        for c in cmds:
            print(" ".join(c))
        return 0
    os.makedirs(a.out_dir, exist_ok=True)
    for name in ("result.csv", "result.jsonl"):
        out = os.path.join(a.out_dir, name)
        if not a.keep and os.path.exists(out):
            os.remove(out)
    failed = []
    for c in cmds:
        print(" ".join(c[1:]), flush=True)
        if not os.path.exists(c[1]):
            failed.append((c, "missing file"))
            continue
        try:
            r = subprocess.run(c, timeout=a.timeout, cwd=a.out_dir)
            if r.returncode != 0:
                failed.append((c, f"exit {r.returncode}"))
        except subprocess.TimeoutExpired:
            failed.append((c, "timeout"))
    print("finish", flush=True)
    for c, why in failed:
        print(f"failed: {' '.join(c[1:])}: {why}", file=sys.stderr)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
