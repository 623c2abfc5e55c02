"""Autotuner driver: searches the launch parameters of every tunable operator
at the size classes the benchmark configurations use, on the device, and
writes the tune file libqo loads at init.

    python -m query_on_opencl_fpga_amd.autotune [--out PATH] [--json PATH] [--reps 5]

Each (operator, size class) runs qo_autotune: the operator's built-in default
and every candidate are timed on a synthetic workload of that size (median of
`reps` runs), and the fastest is kept.  The table is then saved with
qo_tune_save (default: tune/<arch>.tune beside libqo.so, where qo_init finds
it).  The JSON summary lists default and tuned times per entry.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
from pathlib import Path

# (operator, name, size classes): classes of n for streaming operators, of the
# mean build side |R| for the batched join.
PLAN = [
    (0, "filter_agg", [20, 24, 28]),
    (3, "groupby direct (1024 groups)", [20, 24, 28]),
    (5, "partitioned hash join build (LDS slice KiB)", [20, 22, 24]),
    (6, "groupby partitioned (2^20 groups)", [20, 24, 28]),
    (8, "batched join (tiles per probe ticket, build tile)", [12, 14, 16, 18, 20, 22]),
    (9, "batched join table load %", [12, 16, 20, 22]),
]


def main(argv=None) -> int:
    from . import _lib, device

    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=None, help="tune file to write (default: tune/<arch>.tune beside libqo.so)")
    ap.add_argument("--json", default=None, help="also write a JSON summary here")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--ops", default=None, help="comma-separated operator ids (default: all in the plan)")
    ap.add_argument("--max-class", type=int, default=40, help="skip classes above this (quick runs)")
    args = ap.parse_args(argv)
    lib = _lib.load()
    _lib.call("qo_init", 0)
    arch = device.device_info()["arch"].split(":")[0]
    out = Path(args.out) if args.out else _lib.LIB_PATH.parent / "tune" / f"{arch}.tune"
    ops = {int(x) for x in args.ops.split(",")} if args.ops else None
    _lib.call("qo_tune_reset")  # search from the built-in defaults, not from a loaded file
    rows = []
    for op, name, classes in PLAN:
        if ops is not None and op not in ops:
            continue
        for cls in classes:
            if cls > args.max_class:
                continue
            b, it, pa = C.c_int(), C.c_int(), C.c_int()
            best, dflt = C.c_double(), C.c_double()
            _lib.call("qo_autotune", op, cls, args.reps, C.byref(b), C.byref(it), C.byref(pa), C.byref(best),
                      C.byref(dflt))
            row = {"op": op, "name": name, "size_class": cls, "block": b.value, "items": it.value,
                   "param": pa.value, "tuned_ms": best.value, "default_ms": dflt.value,
                   "gain": dflt.value / best.value if best.value > 0 else None}
            rows.append(row)
            print(f"[autotune] op {op} ({name}) class {cls}: default {dflt.value:.3f} ms -> "
                  f"{best.value:.3f} ms with (block {b.value}, items {it.value}, param {pa.value})", flush=True)
    out.parent.mkdir(parents=True, exist_ok=True)
    _lib.call("qo_tune_save", str(out).encode())
    print(f"[autotune] wrote {out}", flush=True)
    if args.json:
        Path(args.json).write_text(json.dumps({"arch": arch, "reps": args.reps, "entries": rows}, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
