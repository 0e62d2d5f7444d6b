"""Fused runs of 27 and 30 qubits, whose plans store every pass but the last out
of place in the order that makes the next pass's tile contiguous, against the
one-sweep-per-gate path (qse_run_sweeps: no planner, no tiles; pinned to the
oracle by tests/test_gpu_parity.py). Random layered circuits and random mixed
lists of every gate kind (SWAPs, multi-controlled phases and the generic 2x2
included), complex128 to 1e-12 and complex64 to 1e-5 (max abs over the whole
state, compared in chunks on the host). Each list runs twice: the first fused
run interprets the plan (generic kernel), the second runs the run-time
specialised kernels compiled for it."""
import numpy as np
import pytest

from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu

CHUNK = 1 << 24


def max_abs_diff(a, b, n):
    worst = 0.0
    for first in range(0, 1 << n, CHUNK):
        count = min(CHUNK, (1 << n) - first)
        worst = max(worst, float(np.max(np.abs(a.download_range(first, count) - b.download_range(first, count)))))
    return worst


@pytest.mark.parametrize("n,prec,kind,tol", [
    (27, capi.C128, "layered", 1e-12),
    (27, capi.C128, "mixed", 1e-12),
    (28, capi.C64, "layered", 1e-5),
    (30, capi.C128, "layered", 1e-12),
])
def test_contiguous_tile_plans_equal_sweeps(ctx, n, prec, kind, tol):
    print(f"contiguous-tile plan vs sweeps n={n} prec={prec} {kind}", flush=True)
    ops = (G.random_layered(n, 6, seed=n) if kind == "layered" else G.random_mixed(n, 200, seed=n)).ops()
    info = ctx.plan_info(n, prec, ops)
    assert info["passes"] >= 2
    fused, swept = ctx.state(n, prec), ctx.state(n, prec)
    try:
        fused.set_random(n)
        swept.set_random(n)
        for _ in range(2):
            fused.run(ops)
            swept.run(ops, fused=False)
        ctx.sync()
        assert max_abs_diff(fused, swept, n) < tol
    finally:
        fused.close()
        swept.close()
