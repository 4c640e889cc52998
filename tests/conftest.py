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
    """Per-field parity of this session: the worst relative error of each field and how
    many comparisons passed by the 1e-5 relative bound or only by the FP32-mirror bound.
    With PARITY_REPORT set, every comparison is also written there as JSON."""
    try:
        import parity
    except ImportError:
        return
    recs = parity.RECORDS
    if not recs:
        return
    tr = terminalreporter
    tr.write_sep("-", "parity: |gpu - fp64| / max|fp64| per field (one-substep bar)")
    by = {}
    for test, field, e_gpu, e_32, scale, clause in recs:
        r = by.setdefault(field, dict(n=0, rel=0, fp32=0, ulp=0, worst=0.0, worst32=0.0, where=""))
        r["n"] += 1
        if clause in ("rel", "fp32", "ulp"):
            r[clause] += 1
        rel = e_gpu / scale if scale > 0 else 0.0
        if rel >= r["worst"]:
            r["worst"], r["worst32"], r["where"] = rel, (e_32 / scale if scale > 0 else 0.0), test
    for field, r in sorted(by.items()):
        tr.write_line(f"  {field:>6}: {r['n']:4d} comparisons, {r['rel']:4d} within 1e-5, {r['fp32']:4d} by the "
                      f"FP32 mirror bound, {r['ulp']:4d} by C's FP32 resolution; worst {r['worst']:.2e} "
                      f"(mirror {r['worst32']:.2e}) in {r['where']}")
    path = os.environ.get("PARITY_REPORT")
    if path:
        import json
        with open(path, "w") as f:
            json.dump([dict(test=t, field=fl, err_gpu=a, err_fp32=b, scale=c, clause=d)
                       for t, fl, a, b, c, d in recs], f, indent=0)
