"""Program ring (device copies of fused programs, overwritten in stream order):
thousands of distinct programs in one call wrap the ring several times, and
every result still matches the oracle. Also the structure-keyed kernels that
start compiling in the background mid-call."""
import math

import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


def test_many_distinct_programs_wrap_the_ring(ctx):
    n = 11
    g = G.GateList(n)
    p = 0
    for layer in range(6):
        for q in range(n):
            g.rotate_y(0.0, q, param=p, scale=1.0)
            g.rotate_z(0.0, q, param=p + 1, scale=1.0)
            p += 2
        for q in range(n - 1):
            g.cnot(q, q + 1)
    ops = g.ops()
    nsets = 3000
    params = np.random.default_rng(21).random((nsets, p)) * 2 * math.pi
    masks, w = G.zsum_observable(n)
    E = ctx.batch_expect(n, capi.C128, ops, params, masks, w)
    for b in list(range(0, nsets, 97)) + [nsets - 1]:
        psi = O.run_ops(O.basis(n), n, ops, params[b])
        assert abs(E[b] - O.expect_zstrings(psi, n, masks, w)) < 1e-11, b
