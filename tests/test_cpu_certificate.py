"""Certificates (include/pramaan_cert.h, pramaan/certificate.hpp) on the CPU.

V1 must be the reference's PRAMAAN_CERTIFICATE_V1 byte for byte: checked
against the reference's own writer (src/certificate/certificate.cpp:23-153)
and verifier (src/certificate/verifier_main.cpp), compiled from the reference
sources into oracle/_ref, on afiro and adlittle with the reference simplex's
x and with perturbed x.  The tamper cases follow the reference's
test_certificate.cpp:609-750 (x, objective, residual, dimension,
fingerprint) plus the reader's rejections; every verdict must agree with the
reference verifier's.

V2 (dual y) has no reference counterpart; its dual measures are pinned by
hand-solved LPs and by the closed-form optimum of the generator (x*, -y*),
where the dual residual, complementarity and the gap are zero up to
rounding.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
REF = ROOT / "oracle" / "_ref"
ref_missing = pytest.mark.skipif(not (REF / "ref_driver").exists() or not (REF / "pramaan_verify").exists(),
                                 reason="reference CPU programs (oracle/_ref) not built")


@pytest.fixture(scope="module")
def P():
    from pramaan_amd import _capi
    _capi.cxx()
    return _capi


def read(name):
    from pramaan_amd import mps
    return mps.read_mps(GOLDEN / name)


def ref_simplex_x(mps_name, tmp_path):
    out = tmp_path / "x.txt"
    r = subprocess.run([str(REF / "ref_driver"), "simplex", str(GOLDEN / mps_name), str(out)], capture_output=True,
                       text=True, check=True)
    assert r.stdout.split()[0] == "0"  # kOptimal
    return np.loadtxt(out)


def ref_cert(mps_name, x, status, tol, path, tmp_path):
    xf = tmp_path / "xin.txt"
    np.savetxt(xf, x, fmt="%.17g")
    subprocess.run([str(REF / "ref_driver"), "cert", str(GOLDEN / mps_name), str(xf), status, repr(tol), str(path)],
                   check=True, capture_output=True)


def ref_verify(mps_name, path) -> int:
    r = subprocess.run([str(REF / "pramaan_verify"), str(GOLDEN / mps_name), str(path)], capture_output=True,
                       text=True)
    assert ("CERTIFICATE VALID" in r.stdout) == (r.returncode == 0)
    return r.returncode


@ref_missing
@pytest.mark.parametrize("name", ["afiro.mps", "adlittle.mps"])
@pytest.mark.parametrize("perturb", [0.0, 1e-3, 5.0])
def test_v1_bytes_match_reference_writer(P, tmp_path, name, perturb):
    lp = read(name)
    x = ref_simplex_x(name, tmp_path)
    x = x + perturb * np.random.default_rng(7).standard_normal(len(x))
    for status, tol in [("OPTIMAL", 1e-6), ("ITERATION_LIMIT", 3e-5)]:
        mine, ref = tmp_path / "mine.cert", tmp_path / "ref.cert"
        P.cert_write(lp, x, mine, status=status, tolerance=tol)
        ref_cert(name, x, status, tol, ref, tmp_path)
        assert mine.read_bytes() == ref.read_bytes()


def tamper(text: str, key: str, value: str) -> str:
    out = []
    for line in text.splitlines(keepends=True):
        out.append(f"{key} {value}\n" if line.startswith(key + " ") else line)
    return "".join(out)


def tamper_x0(text: str, delta: float) -> str:
    lines = text.splitlines(keepends=True)
    i = lines.index("X\n") + 1
    lines[i] = repr(float(lines[i]) + delta) + "\n"
    return "".join(lines)


CASES = {
    "valid": lambda t: t,
    "x0+50": lambda t: tamper_x0(t, 50.0),
    "objective+100": lambda t: tamper(t, "OBJECTIVE", "12345.5"),
    "primal_residual=50": lambda t: tamper(t, "PRIMAL_RESIDUAL", "50"),
    "cols=999": lambda t: tamper(t, "COLS", "999"),
    "fingerprint=0": lambda t: tamper(t, "MODEL_FINGERPRINT", "0"),
    "status": lambda t: tamper(t, "STATUS", "ITERATION_LIMIT"),
    "tolerance=1e-3": lambda t: tamper(t, "TOLERANCE", "0.001"),
    "tolerance=0": lambda t: tamper(t, "TOLERANCE", "0"),
    "dual_residual given": lambda t: tamper(t, "DUAL_RESIDUAL", "0"),
    "gap given": lambda t: tamper(t, "OBJECTIVE_BOUND_GAP", "0"),
    "not_computed": lambda t: tamper(t, "DUAL_RESIDUAL", "NOT_COMPUTED"),
    "nan x": lambda t: t.replace("X\n", "X\nnan\n", 1),
    "garbage": lambda t: tamper(t, "ROWS", "27abc"),
    "trailing": lambda t: t + "more\n",
    "blank trailing": lambda t: t + "\n  \n",
    "crlf": lambda t: t.replace("\n", "\r\n"),
    "truncated": lambda t: t[: t.index("END_X")],
    "format": lambda t: t.replace("PRAMAAN_CERTIFICATE_V1", "PRAMAAN_CERTIFICATE_V0"),
    "bound violation": lambda t: tamper(t, "BOUND_VIOLATION", "1e-3"),
}


@ref_missing
@pytest.mark.parametrize("case", sorted(CASES))
def test_verdicts_match_reference_verifier(P, tmp_path, case):
    lp = read("afiro.mps")
    x = ref_simplex_x("afiro.mps", tmp_path)
    good = tmp_path / "good.cert"
    P.cert_write(lp, x, good, tolerance=1e-6)
    bad = tmp_path / "bad.cert"
    bad.write_bytes(CASES[case](good.read_text()).encode())
    rc, msgs = P.cert_verify(lp, bad)
    ref_rc = ref_verify("afiro.mps", bad)
    assert (rc == 0) == (ref_rc == 0), (case, rc, msgs)
    expect_valid = case in ("valid", "not_computed", "blank trailing", "crlf")
    assert (rc == 0) == expect_valid, (case, msgs)


def test_v1_rejects_infeasible_point(P, tmp_path):
    # min x1 + x2, x1 + x2 >= 2, 0 <= x <= 10: x = (0.5, 0.5) violates the row by 1.
    lp = P.LP.from_dense([[1, 1]], [1, 1], [2], [1e30], [0, 0], [10, 10])
    f = tmp_path / "c.cert"
    P.cert_write(lp, [0.5, 0.5], f, tolerance=1e-6)
    rc, msgs = P.cert_verify(lp, f)
    assert rc == 1 and any("Primal infeasible" in m for m in msgs)
    P.cert_write(lp, [1.0, 1.0], f, tolerance=1e-6)
    assert P.cert_verify(lp, f) == (0, [])
    with pytest.raises(ValueError):
        P.cert_write(lp, [1.0], f)


def parse(path):
    vals = {}
    for line in Path(path).read_text().splitlines():
        k, _, v = line.partition(" ")
        vals[k] = v
    return vals


def test_v2_hand_solved_duals(P, tmp_path):
    """min x s.t. x >= 1 (row), 0 <= x <= 10.  Optimum x = 1; the reference-
    sign dual makes the reduced cost 1 + y vanish, so y = -1.

    y = -1:   all dual measures 0.
    y = -0.5: reduced cost 0.5 sits on x's lower bound 0 (allowed sign), dual
              objective 0.5*1 + 0.5*0 = 0.5, gap 0.5, complementarity
              0.5*(x - 0) = 0.5, dual residual 0.
    y = +1:   row multiplier -1 on a row without upper bound: residual 1,
              and that multiplier is dropped; reduced cost 2 on the lower
              bound: complementarity 2*(1 - 0) = 2, dual objective 2*0 = 0,
              gap |1 - 0| = 1.
    """
    lp = P.LP.from_dense([[1.0]], [1.0], [1.0], [1e30], [0.0], [10.0])
    f = tmp_path / "v2.cert"
    for y, res, comp, gap in [(-1.0, 0.0, 0.0, 0.0), (-0.5, 0.0, 0.5, 0.5), (1.0, 1.0, 2.0, 1.0)]:
        P.cert_write(lp, [1.0], f, y=[y], tolerance=1e-6)
        v = parse(f)
        assert v["PRAMAAN_CERTIFICATE_V2"] == ""
        assert float(v["DUAL_RESIDUAL"]) == pytest.approx(res, abs=1e-15)
        assert float(v["COMPLEMENTARITY_RESIDUAL"]) == pytest.approx(comp, abs=1e-15)
        assert float(v["OBJECTIVE_BOUND_GAP"]) == pytest.approx(gap, abs=1e-15)
        rc, msgs = P.cert_verify(lp, f)
        assert (rc == 0) == (y == -1.0), msgs
    # Maximisation: max x s.t. x <= 3 (row), 0 <= x <= 10.  In minimisation
    # form min -x; reduced cost -1 + y = 0 gives y = 1, row multiplier -1 on
    # the upper bound; the gap is |-3 - (-1*3)| = 0.
    lp = P.LP.from_dense([[1.0]], [1.0], [-1e30], [3.0], [0.0], [10.0], sense=1)
    P.cert_write(lp, [3.0], f, y=[1.0], tolerance=1e-6)
    v = parse(f)
    assert float(v["OBJECTIVE"]) == 3.0
    assert float(v["OBJECTIVE_BOUND_GAP"]) == 0.0 and float(v["DUAL_RESIDUAL"]) == 0.0
    assert P.cert_verify(lp, f) == (0, [])


@pytest.mark.parametrize("which,seed", [(0, 1), (None, 5)])
def test_v2_closed_form_optimum(P, tmp_path, which, seed):
    lp = P.generate(which, seed=seed) if which is not None else P.generate(None, seed=seed, m=300, n=750,
                                                                             eq_frac=0.6)
    f = tmp_path / "v2.cert"
    P.cert_write(lp, lp.x_star, f, y=-lp.y_star, tolerance=1e-6)
    v = parse(f)
    scale = 1 + abs(lp.p_star)
    assert float(v["OBJECTIVE"]) == pytest.approx(lp.p_star, rel=1e-12)
    assert float(v["DUAL_RESIDUAL"]) <= 1e-12 * scale
    assert float(v["COMPLEMENTARITY_RESIDUAL"]) <= 1e-12 * scale
    assert float(v["OBJECTIVE_BOUND_GAP"]) <= 1e-10 * scale
    assert P.cert_verify(lp, f) == (0, [])
    # The textbook sign of y is not the reference's: the checker must refuse it.
    P.cert_write(lp, lp.x_star, f, y=lp.y_star, tolerance=1e-6)
    rc, msgs = P.cert_verify(lp, f)
    assert rc == 1 and any("Dual" in m or "Complementarity" in m for m in msgs), msgs
    # A written dual field that disagrees with the recomputed one is refused.
    P.cert_write(lp, lp.x_star, f, y=-lp.y_star, tolerance=1e-6)
    f.write_text(tamper(f.read_text(), "OBJECTIVE_BOUND_GAP", "0.5"))
    rc, msgs = P.cert_verify(lp, f)
    assert rc == 1 and any("gap mismatch" in m for m in msgs)
    # A missing Y section is a parse error.
    P.cert_write(lp, lp.x_star, f, y=-lp.y_star, tolerance=1e-6)
    t = f.read_text()
    f.write_text(t[: t.index("Y\n")] + "END_CERTIFICATE\n")
    assert P.cert_verify(lp, f)[0] == 2
