import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def pipeline():
    from nexus_amd import mpm
    p = mpm.MpmPipeline(0)
    yield p
    p.close()


def pytest_terminal_summary(terminalreporter):
    """Per-field parity of this session: how many comparisons passed by each clause of the
    one-substep bar and, per clause, the worst margin: the largest |gpu - fp64| / max|fp64|
    among those within 1e-5, the largest |gpu - fp64| / |fp32 mirror - fp64| among those
    passing by the mirror bound (2 allowed), the largest |gpu - fp64| / resolution among
    those passing by C's FP32 resolution (4 allowed).  With PARITY_REPORT set, every
    comparison is also written there as JSON."""
    try:
        import parity
    except ImportError:
        return
    recs = parity.RECORDS
    if not recs:
        return
    tr = terminalreporter
    tr.write_sep("-", "parity per field: clause counts and worst margin per clause (one-substep bar)")
    by = {}
    for test, field, e_gpu, e_32, scale, clause, res in recs:
        r = by.setdefault(field, {"n": 0, "rel": [0, 0.0, ""], "fp32": [0, 0.0, ""], "ulp": [0, 0.0, ""]})
        r["n"] += 1
        if clause == "rel":
            m = e_gpu / scale if scale > 0 else 0.0
        elif clause == "fp32":
            m = e_gpu / e_32 if e_32 > 0 else 0.0
        elif clause == "ulp":
            m = e_gpu / res if res > 0 else 0.0
        else:
            continue
        c = r[clause]
        c[0] += 1
        if m >= c[1]:
            c[1], c[2] = m, test.split("::")[-1]
    for field, r in sorted(by.items()):
        parts = [f"{field:>22}: {r['n']:3d}"]
        for clause, label in (("rel", "within 1e-5, worst rel"), ("fp32", "by mirror, worst gpu/mirror"),
                              ("ulp", "by C resolution, worst gpu/res")):
            n, m, where = r[clause]
            if n:
                parts.append(f"{n} {label} {m:.2e} ({where})")
        tr.write_line("; ".join(parts))
    path = os.environ.get("PARITY_REPORT")
    if path:
        import json
        with open(path, "w") as f:
            json.dump([dict(test=t, field=fl, err_gpu=a, err_fp32=b, scale=c, clause=d, c_resolution=e)
                       for t, fl, a, b, c, d, e in recs], f, indent=0)
