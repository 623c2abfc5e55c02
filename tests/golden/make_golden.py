"""Writes tests/golden/brute_force.json: small seeded inputs with outputs from
straightforward pure-Python implementations of each operator's definition
(independent of oracle/qo_oracle.c, which the fixtures then pin).

    python tests/golden/make_golden.py

The definitions restated here are the closed forms the reference's own
validators check (host/host/Validate.cpp: validateScan :4-37, validateFilter
:317-370, validateSort :470-487, validateGroupBy :257-316) plus the
reference's aggregate operator semantics (kernel/aggregation/1/
group_by_cu_op_dp.cl:547-716) and the hash-join match definition
(kernel/hj_filter_sort/1/hj_filter_sort_op.cl:62-109), and the results the
split / scatter / gather / projection / reduce drivers and validators fix
(host/host/testSplit.cpp:156-209, testScatter.cpp:61-130, Validate.cpp:39-71,
372-468, 510-527).
"""
from __future__ import annotations

import json
import random
from pathlib import Path

OUT = Path(__file__).resolve().parent / "brute_force.json"
M32 = 1 << 32


def s32(v: int) -> int:
    v &= M32 - 1
    return v - M32 if v >= 1 << 31 else v


def scan(xs):
    out, acc = [], 0
    for x in xs:
        out.append(s32(acc))
        acc += x
    return out


def filt(recs, lo, hi):
    return [r for r in recs if lo <= s32(r[1]) <= hi]


def sort_stable(recs):
    return sorted(recs, key=lambda r: r[1])  # Python's sort is stable; y compared unsigned


def startpos(sorted_recs):
    return [i for i in range(len(sorted_recs)) if i == 0 or sorted_recs[i][1] != sorted_recs[i - 1][1]]


def paragg(sorted_recs, sp, op):
    out = []
    for g, s in enumerate(sp):
        e = sp[g + 1] if g + 1 < len(sp) else len(sorted_recs)
        vals = [s32(r[0]) for r in sorted_recs[s:e]]
        if op == 1:
            out.append(max([0] + vals))
        elif op == 2:
            out.append(min([1 << 30] + vals))
        else:
            t = s32(sum(vals))
            if op == 3:
                c = e - s
                q = abs(t) // c
                out.append(q if t >= 0 else -q)
            else:
                out.append(t)
    return out


def groupby(keys, vals):
    d = {}
    for k, v in zip(keys, vals):
        c, s, mn, mx = d.get(k, (0, 0, 1 << 31, -(1 << 31)))
        d[k] = (c + 1, s + v, min(mn, v), max(mx, v))
    ks = sorted(d)
    return [[k, *d[k]] for k in ks]


def join(rk, rp, sk, sp, sf, flo, fhi):
    rows = []
    for j in range(len(sk)):
        if sf is not None and not (flo <= sf[j] <= fhi):
            continue
        for i in range(len(rk)):
            if rk[i] == sk[j]:
                rows.append([sk[j], rp[i], sp[j]])
    return rows


def split(recs, parts):
    """Bucket by y % parts, keeping input order inside each bucket (validateSplit,
    Validate.cpp:510-527, checks the order of the bucket ids); bounds per bucket."""
    buckets = [[] for _ in range(parts)]
    for r in recs:
        buckets[r[1] % parts].append(r)
    out, bounds, pos = [], [], 0
    for b in buckets:
        bounds.append([pos, pos + len(b)])
        out.extend(b)
        pos += len(b)
    return out, bounds


def gather(recs, loc, s_init):
    return [list(recs[l]) if 0 <= l < len(recs) else list(s) for l, s in zip(loc, s_init)]


def scatter(recs, loc, s_init):
    s = [list(x) for x in s_init]
    for r, l in zip(recs, loc):
        if 0 <= l < len(recs):
            s[l] = list(r)
    return s


def project(recs, proj):
    return [[p[0], recs[p[0]][1] if p[0] < len(recs) else p[1]] for p in proj]


def reduce_(recs, op):
    """validateReduce (Validate.cpp:372-468) on uint .y."""
    ys = [r[1] for r in recs]
    if not ys:
        return 0
    if op == 1:
        return max(ys)
    if op == 2:
        return min(ys)
    t = sum(ys) % M32
    return t // len(ys) if op == 3 else t


def main() -> None:
    rnd = random.Random(20240611)
    cases = {"scan": [], "filter": [], "sort": [], "groupby_sorted": [], "groupby_hash": [], "join": []}
    for n in (0, 1, 7, 256, 300, 513):
        xs = [rnd.randrange(-(1 << 31), 1 << 31) for _ in range(n)]
        cases["scan"].append({"in": xs, "out": scan(xs)})
    for n, lo, hi in ((0, 0, 5), (1, 0, 0), (200, 10, 60), (333, -50, 50), (500, -(1 << 31), (1 << 31) - 1)):
        recs = [[rnd.randrange(M32), rnd.randrange(-100, 100) & (M32 - 1)] for _ in range(n)]
        cases["filter"].append({"in": recs, "lo": lo, "hi": hi, "out": filt(recs, lo, hi)})
    for n, kmax in ((1, 10), (50, 5), (300, 1 << 32), (400, 17)):
        recs = [[i, rnd.randrange(kmax)] for i in range(n)]
        cases["sort"].append({"in": recs, "out": sort_stable(recs)})
    for n, kmax in ((5, 3), (300, 20), (400, 400)):
        recs = sort_stable([[rnd.randrange(-(1 << 31), 1 << 31) & (M32 - 1), rnd.randrange(kmax)] for _ in range(n)])
        sp = startpos(recs)
        cases["groupby_sorted"].append({"in": recs, "startPos": sp, "agg": {str(op): paragg(recs, sp, op)
                                                                            for op in range(4)}})
    for n, kmax in ((0, 1), (10, 3), (400, 37), (500, 1000)):
        keys = [rnd.randrange(-kmax, kmax) for _ in range(n)]
        vals = [rnd.randrange(-(1 << 31), 1 << 31) for _ in range(n)]
        cases["groupby_hash"].append({"key": keys, "val": vals, "out": groupby(keys, vals)})
    for nr, ns, dup in ((0, 5, False), (20, 100, False), (40, 200, True), (64, 300, True)):
        if dup:
            rk = [rnd.randrange(-10, 10) for _ in range(nr)]
        else:
            rk = rnd.sample(range(-1000, 1000), nr)
        rp = [rnd.randrange(-(1 << 31), 1 << 31) for _ in range(nr)]
        sk = [rnd.randrange(-20, 20) if dup else rnd.randrange(-1000, 1000) for _ in range(ns)]
        sp = [rnd.randrange(-(1 << 31), 1 << 31) for _ in range(ns)]
        sf = [rnd.randrange(100) for _ in range(ns)]
        rows = join(rk, rp, sk, sp, sf, 0, 49)
        cases["join"].append({"rkey": rk, "rpay": rp, "skey": sk, "spay": sp, "sfilt": sf, "flo": 0, "fhi": 49,
                              "rows": rows,
                              "sorted": sorted(rows, key=lambda r: (r[0], r[2]))})
    # Split / gather / scatter / projection / reduce (drawn after the cases
    # above, so their inputs are unchanged).
    cases.update({"split": [], "gather": [], "scatter": [], "project": [], "reduce": []})
    for n, parts, kmax in ((0, 3, 10), (1, 1, 5), (37, 4, 1 << 32), (300, 32, 1 << 32), (500, 300, 1 << 32),
                           (200, 7, 7)):
        recs = [[i, rnd.randrange(kmax)] for i in range(n)]
        out, bounds = split(recs, parts)
        cases["split"].append({"in": recs, "parts": parts, "out": out, "bound": bounds})
    for rl, sl in ((0, 4), (5, 0), (40, 60), (300, 300)):
        recs = [[rnd.randrange(M32), rnd.randrange(M32)] for _ in range(rl)]
        loc = [rnd.randrange(-3, rl + 3) for _ in range(sl)]
        s_init = [[rnd.randrange(M32), rnd.randrange(M32)] for _ in range(sl)]
        cases["gather"].append({"R": recs, "loc": loc, "S0": s_init, "S": gather(recs, loc, s_init)})
    for rl in (0, 1, 50, 257):
        recs = [[rnd.randrange(M32), rnd.randrange(M32)] for _ in range(rl)]
        loc = [rnd.randrange(-2, rl + 2) for _ in range(rl)]
        s_init = [[rnd.randrange(M32), rnd.randrange(M32)] for _ in range(rl)]
        cases["scatter"].append({"R": recs, "loc": loc, "S0": s_init, "S": scatter(recs, loc, s_init)})
    for rl, pl in ((0, 3), (10, 20), (300, 100)):
        recs = [[rnd.randrange(M32), rnd.randrange(M32)] for _ in range(rl)]
        proj = [[rnd.randrange(rl + 4), rnd.randrange(M32)] for _ in range(pl)]
        cases["project"].append({"R": recs, "proj": proj, "out": project(recs, proj)})
    for n, kmax in ((0, 5), (1, M32), (9, 100), (400, M32), (513, 1 << 31)):
        recs = [[rnd.randrange(M32), rnd.randrange(kmax)] for _ in range(n)]
        cases["reduce"].append({"in": recs, "out": [reduce_(recs, op) for op in range(4)]})
    OUT.write_text(json.dumps(cases))
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
