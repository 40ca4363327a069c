"""Free-format MPS reader producing the engine's LP (ModelIR layout).

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
