"""Both Pauli-noise modes through the Circuit class on the GPU, against the
oracle: deferred (the reference's frame after the last gate, Circuit.cpp:64-120,
Noise.cpp:34-44) and interleaved (each drawn flip right after its gate, no
counterpart in the reference). The Circuit's own Noise (seed 42) makes the
draws; the oracle replays the same trajectory's gate list or frame."""
import numpy as np
import pytest

import oracle_lib as O
from quantumsimevo_amd import Circuit, capi
from quantumsimevo_amd import circuits as C
from quantumsimevo_amd import _quantumsimevo as Q

NOISY = (capi.X, capi.Y, capi.Z, capi.H, capi.CNOT, capi.PHASE, capi.RX, capi.RY, capi.RZ)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["deferred", "interleaved"])
def test_circuit_noise_modes_match_oracle(tmp_path, mode):
    n, rate = 10, 0.05
    ops = C.random_layered(n, 8, seed=4).ops()
    path = tmp_path / "c.qsegates"
    C.save_gates(path, n, ops)
    qc = Circuit(n, True, "c128", noise_mode=mode)
    assert qc.noise_mode == mode
    for k in NOISY:
        qc.set_noise_rate(int(k), rate)
    qc.load_gates(str(path))
    qc.execute()
    psi = np.asarray(qc.state())
    print(f"mode={mode} n={n} stats={qc.noise_stats()}", flush=True)
    if mode == "interleaved":
        traj = C.interleaved_trajectories(n, ops, [42], rate, 1.0)[0]
        assert len(traj) > len(ops)  # the draws hit at this rate
        ref = O.run_ops(O.basis(n), n, traj)
    else:
        x, z = Q.noise_frames(n, ops["kind"], ops["target"], np.array([42], dtype=np.uint32), rate, 1.0)
        frame = C.GateList(n)
        for q in range(n):  # X_q then Z_q, ascending q (Noise.cpp:34-44)
            if (int(x[0]) >> q) & 1:
                frame.pauli_x(q)
            if (int(z[0]) >> q) & 1:
                frame.pauli_z(q)
        ref = O.run_ops(O.run_ops(O.basis(n), n, ops), n, frame.ops())
    assert np.max(np.abs(psi - ref)) < 1e-12
    # the same draws in both modes: the counts and the frame the deferred mode applies
    st = qc.noise_stats()
    x, z = Q.noise_frames(n, ops["kind"], ops["target"], np.array([42], dtype=np.uint32), rate, 1.0)
    assert st["x_mask"] == int(x[0]) and st["z_mask"] == int(z[0])


@pytest.mark.gpu
def test_interleaved_trajectories_sample_like_the_oracle(ctx):
    # interleaved trajectories run through the C ABI as their own gate lists; sampled outcomes
    # equal the oracle's sequential-CDF samples of the oracle's state (states within 1e-12)
    n, rate = 9, 0.03
    ops = C.random_layered(n, 6, seed=12).ops()
    trajs = C.interleaved_trajectories(n, ops, np.arange(8, dtype=np.uint32), rate, 1.0)
    u = O.uniforms(3, 400)
    st = ctx.state(n, capi.C128)
    try:
        for t in trajs:
            st.set_basis(0)
            st.run(t)
            psi = st.download()
            ref = O.run_ops(O.basis(n), n, t)
            assert np.max(np.abs(psi - ref)) < 1e-12
            out = st.sample(u)
            assert np.array_equal(np.asarray(out, dtype=np.int64), np.asarray(O.sample(psi, n, u), dtype=np.int64))
    finally:
        st.close()
