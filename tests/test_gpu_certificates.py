"""GPU solutions certified end to end.

The certificate is the reference's PRAMAAN_CERTIFICATE_V1 (primal only,
src/certificate/certificate.cpp:23-153) or this project's V2, which adds
the FP64 dual and its residuals.  Solutions come from the GPU engines and
are judged by the checkers, never by the solver's own numbers:
  * afiro and adlittle (the reference's Netlib files), solved by the
    FP64-refined PDHG on the single-LP engine under the reference's own
    termination test (pv, dv <= the tolerance): a V1 certificate accepted
    by the reference's pramaan-verify built from its sources
    (src/certificate/verifier_main.cpp) and by this project's checker, and
    a V2 certificate of the same point accepted by this project's checker;
  * rolling-horizon windows solved in one batch by the batch engine: V2
    certificates accepted by this project's checker, whose dual-residual,
    complementarity and gap checks are the FP64 KKT judgement of each
    solution.
"""
import json
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
VERIFY = ROOT / "oracle" / "_ref" / "pramaan_verify"


@pytest.mark.parametrize("name", ["afiro", "adlittle"])
def test_netlib_gpu_solution_certified(capi, tmp_path, name):
    from pramaan_amd.mps import read_mps
    lp = read_mps(GOLDEN / f"{name}.mps")
    tol = 1e-6
    r = capi.solve(lp, algorithm="pdlp", termination="reference", tolerance=tol, max_iterations=400000,
                   check_frequency=64)
    print(f"{name}: {r.status} it={r.iterations} pv={r.primal_residual:.2e} dv={r.dual_residual:.2e} "
          f"rel_kkt={r.rel_kkt:.2e} obj={r.objective:.12g}", flush=True)
    assert r.status == "OPTIMAL"
    want = json.loads((GOLDEN / "fixtures.json").read_text())["netlib"][name]
    assert abs(r.objective - want) <= 1e-6 * (1 + abs(want))
    v1, v2 = tmp_path / "v1.cert", tmp_path / "v2.cert"
    capi.cert_write(lp, r.x, v1, tolerance=tol)
    capi.cert_write(lp, r.x, v2, y=r.y, tolerance=tol)
    assert capi.cert_verify(lp, v1) == (0, []), capi.cert_verify(lp, v1)
    rc, msgs = capi.cert_verify(lp, v2)
    assert rc == 0, msgs
    if not VERIFY.exists():
        pytest.skip("the reference's pramaan-verify was not built (needs the reference sources at build time)")
    p = subprocess.run([str(VERIFY), str(GOLDEN / f"{name}.mps"), str(v1)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr


def test_batch_window_solutions_certified(capi, tmp_path):
    rh = capi.RollingHorizon()
    lps = [rh.window(w) for w in (0, 57, 123, 250, 311, 499)]
    with capi.Batch(lps, tolerance=1e-9, max_iterations=200000, check_frequency=64) as b:
        res, _, _ = b.run(want_vectors=True)
    for w, (lp, r) in enumerate(zip(lps, res)):
        assert r.status == "OPTIMAL", (w, r.status)
        path = tmp_path / f"w{w}.cert"
        capi.cert_write(lp, r.x, path, y=r.y, tolerance=1e-5)
        rc, msgs = capi.cert_verify(lp, path)
        print(f"window {w}: rel_kkt={r.rel_kkt:.2e} verdict {rc} {msgs}", flush=True)
        assert rc == 0, (w, msgs)
