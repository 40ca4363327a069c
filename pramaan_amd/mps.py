"""Free-format MPS reader producing the engine's LP (ModelIR layout), and
the matching writer.

Follows the reference's parse_mps (src/io/mps_parser.cpp:75-490): ROWS
N/L/G/E (the first N row is the objective, later N rows become free rows),
COLUMNS with 'MARKER' INTORG/INTEND, RHS (a value on the objective row is
subtracted from the objective offset, 271-279), BOUNDS LO/UP/FX/FR/MI/PL
(default 0 <= x < +inf, 458-466), RANGES and BV rejected (340-355).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from ._capi import INF, LP


class MpsError(ValueError):
    pass


def read_mps(path: str | Path) -> LP:
    rows: dict[str, tuple[str, int]] = {}   # name -> (sense, index or -1)
    senses: list[str] = []
    obj_name = None
    cols: dict[str, int] = {}
    obj: dict[int, float] = {}
    trip: dict[tuple[int, int], float] = {}
    rhs: dict[int, float] = {}
    bounds: dict[int, list] = {}
    offset = 0.0
    section = None

    def var(name):
        if name not in cols:
            cols[name] = len(cols)
        return cols[name]

    for lineno, raw in enumerate(Path(path).read_text().splitlines(), 1):
        line = raw.rstrip()
        if not line.strip() or line.lstrip().startswith("*"):
            continue
        tok = line.split()
        if not raw[0].isspace():
            head = tok[0].upper()
            if head == "NAME":
                continue
            if head == "ENDATA":
                break
            if head in ("ROWS", "COLUMNS", "RHS", "BOUNDS"):
                section = head
                continue
            if head == "RANGES":
                raise MpsError(f"line {lineno}: RANGES not supported")
            raise MpsError(f"line {lineno}: unknown section {tok[0]}")
        if section == "ROWS":
            sense, name = tok[0].upper(), tok[1]
            if sense == "N" and obj_name is None:
                obj_name = name
                rows[name] = ("N", -1)
            else:
                rows[name] = (sense, len(senses))
                senses.append(sense)
        elif section == "COLUMNS":
            if len(tok) >= 3 and tok[1] == "'MARKER'":
                continue
            if len(tok) < 3 or len(tok) == 4:
                raise MpsError(f"line {lineno}: COLUMNS line has unexpected token count")
            j = var(tok[0])
            for p in range(1, len(tok) - 1, 2):
                rname, val = tok[p], float(tok[p + 1])
                if rname not in rows:
                    raise MpsError(f"line {lineno}: unknown row {rname}")
                idx = rows[rname][1]
                if idx < 0:
                    obj[j] = obj.get(j, 0.0) + val
                else:
                    trip[(idx, j)] = trip.get((idx, j), 0.0) + val
        elif section == "RHS":
            for p in range(1, len(tok) - 1, 2):
                rname, val = tok[p], float(tok[p + 1])
                if rname not in rows:
                    raise MpsError(f"line {lineno}: unknown row {rname}")
                idx = rows[rname][1]
                if idx < 0:
                    offset -= val
                else:
                    rhs[idx] = val
        elif section == "BOUNDS":
            bt, vname = tok[0].upper(), tok[2]
            j = var(vname)
            b = bounds.setdefault(j, [None, None])
            val = float(tok[3]) if len(tok) > 3 else None
            if bt in ("LO", "UP", "FX") and val is None:
                raise MpsError(f"line {lineno}: {bt} bound needs a value")
            if bt == "LO":
                b[0] = val
            elif bt == "UP":
                b[1] = val
            elif bt == "FX":
                b[0] = b[1] = val
            elif bt == "FR":
                b[0], b[1] = -INF, INF
            elif bt == "MI":
                b[0] = -INF
            elif bt == "PL":
                b[1] = INF
            else:
                raise MpsError(f"line {lineno}: {bt} bounds not supported")
    m, n = len(senses), len(cols)
    c = np.zeros(n)
    for j, v in obj.items():
        c[j] = v
    rl, ru = np.zeros(m), np.zeros(m)
    for i, s in enumerate(senses):
        b = rhs.get(i, 0.0)
        rl[i], ru[i] = {"L": (-INF, b), "G": (b, INF), "E": (b, b), "N": (-INF, INF)}[s]
    xl, xu = np.zeros(n), np.full(n, INF)
    for j, (lo, hi) in bounds.items():
        if lo is not None:
            xl[j] = lo
        if hi is not None:
            xu[j] = hi
    keys = sorted(trip)
    rp = np.zeros(m + 1, np.int64)
    for (i, _j) in keys:
        rp[i + 1] += 1
    rp = np.cumsum(rp)
    ci = np.array([j for (_i, j) in keys], np.int32)
    vals = np.array([trip[k] for k in keys], float)
    return LP(c, rp, ci, vals, rl, ru, xl, xu, 0, offset)


def write_mps(lp: LP, path: str | Path, name: str = "LP") -> None:
    """Writes lp as free-format MPS that read_mps and the reference's
    parse_mps (src/io/mps_parser.cpp:75-490) read back: rows E (equal
    bounds), L, G or N (free); bounds LO/UP/FX/FR/MI where they differ from
    the default 0 <= x < +inf; a maximisation is written as the minimisation
    of -c (the format has no sense section there).  A row with two different
    finite bounds would need RANGES, which that parser rejects: MpsError."""
    big = INF * 0.1
    sign = -1.0 if lp.sense == 1 else 1.0
    out = [f"NAME {name}", "ROWS", " N OBJ"]
    kinds = []
    for i in range(lp.m):
        lo, hi = lp.row_lower[i], lp.row_upper[i]
        has_lo, has_hi = lo > -big, hi < big
        if has_lo and has_hi and lo != hi:
            raise MpsError(f"row {i} has two different finite bounds (RANGES)")
        kind = "E" if has_lo and has_hi else ("L" if has_hi else ("G" if has_lo else "N"))
        kinds.append(kind)
        out.append(f" {kind} R{i}")
    out.append("COLUMNS")
    cp = np.zeros(lp.n + 1, np.int64)
    np.add.at(cp, lp.col_idx.astype(np.int64) + 1, 1)
    cp = np.cumsum(cp)
    order = np.argsort(lp.col_idx, kind="stable")
    rows_of = np.repeat(np.arange(lp.m), np.diff(lp.row_ptr))
    for j in range(lp.n):
        if lp.c[j] != 0.0:
            out.append(f" X{j} OBJ {float(sign * lp.c[j])!r}")
        for q in order[cp[j]:cp[j + 1]]:
            out.append(f" X{j} R{rows_of[q]} {float(lp.values[q])!r}")
        if lp.c[j] == 0.0 and cp[j] == cp[j + 1]:
            out.append(f" X{j} OBJ 0.0")
    out.append("RHS")
    for i, kind in enumerate(kinds):
        v = lp.row_upper[i] if kind == "L" else (lp.row_lower[i] if kind in "EG" else 0.0)
        if kind != "N" and v != 0.0:
            out.append(f" RHS R{i} {float(v)!r}")
    if sign * lp.offset != 0.0:
        out.append(f" RHS OBJ {float(-sign * lp.offset)!r}")
    out.append("BOUNDS")
    for j in range(lp.n):
        lo, hi = lp.var_lower[j], lp.var_upper[j]
        has_lo, has_hi = lo > -big, hi < big
        if has_lo and has_hi and lo == hi:
            out.append(f" FX BND X{j} {float(lo)!r}")
            continue
        if not has_lo and not has_hi:
            out.append(f" FR BND X{j}")
            continue
        if not has_lo:
            out.append(f" MI BND X{j}")
        elif lo != 0.0:
            out.append(f" LO BND X{j} {float(lo)!r}")
        if has_hi:
            out.append(f" UP BND X{j} {float(hi)!r}")
    out.append("ENDATA")
    Path(path).write_text("\n".join(out) + "\n")
