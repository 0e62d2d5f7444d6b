"""qse_run_repeat: a gate list applied many times, replayed from a HIP graph
after two eager runs. Every result equals the same number of eager qse_run
calls bit for bit: in-place plans, a plan whose last pass exchanges the state's
buffers (QFT's bit reversal; captured as two runs), an odd count, the
single-gate path below 10 qubits, and graph replays interleaved with eager work
and host waits."""
import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import capi
from quantumsimevo_amd import circuits as G

pytestmark = pytest.mark.gpu


def eager(ctx, n, prec, ops, reps, seed):
    # settle the program's kernels first (under QSE_JIT_AUTO the first run of a
    # program interprets it and the second compiles it): both paths below then
    # run the same kernels, so their results agree bit for bit
    warm = ctx.state(n, prec)
    warm.set_basis(0)
    warm.run(ops)
    warm.run(ops)
    warm.close()
    st = ctx.state(n, prec)
    st.set_random(seed)
    for _ in range(reps):
        st.run(ops)
    out = st.download()
    st.close()
    return out


@pytest.mark.parametrize("n, make, reps", [
    (16, lambda n: G.random_layered(n, 6, seed=3).ops(), 7),
    (14, lambda n: G.qft(n).ops(), 5),
    (15, lambda n: G.qft(n).ops(), 6),
    (8, lambda n: G.random_layered(n, 4, seed=4).ops(), 4),
])
@pytest.mark.parametrize("prec", [capi.C128, capi.C64])
def test_repeat_equals_eager(ctx, n, make, reps, prec):
    ops = make(n)
    want = eager(ctx, n, prec, ops, reps, seed=11)
    st = ctx.state(n, prec)
    st.set_random(11)
    st.run_repeat(ops, reps)
    got = st.download()
    assert np.array_equal(got, want)
    if prec == capi.C128 and n <= 16:
        ref = st.download()  # state after reps runs, checked against the oracle from the same input
        st.set_random(11)
        start = st.download()
        for _ in range(reps):
            start = O.run_ops(start, n, ops)
        assert np.max(np.abs(ref - start)) < 1e-11
    st.close()


def test_replays_interleaved_with_eager_work(ctx):
    n = 14
    ops = G.qft(n).ops()
    st = ctx.state(n, capi.C128)
    st.set_random(5)
    ref = ctx.state(n, capi.C128)
    ref.set_random(5)
    for _ in range(2):  # settle the program's kernels (see eager())
        ref.run(ops)
    ref.set_random(5)
    for rounds in range(3):
        st.run_repeat(ops, 3)
        for _ in range(3):
            ref.run(ops)
        ctx.sync()
        assert np.array_equal(st.download(), ref.download())
        st.run(ops)  # eager work between the replays
        ref.run(ops)
        ctx.sync()
    st.close()
    ref.close()
