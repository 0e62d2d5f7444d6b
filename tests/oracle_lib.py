"""ctypes access to the CPU oracle libraries (test infrastructure).

  oracle/liboracle.so       the restated scalar FP64 oracle (oracle/qse_oracle.cpp)
  oracle/_ref/libqse_ref.so the reference simulator's own CPU backends (built only
                            where the reference tree is present; optional here)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_PATH = ROOT / "oracle" / "liboracle.so"
REF_PATH = ROOT / "oracle" / "_ref" / "libqse_ref.so"

_o = None
_r = None

_vp, _i, _d, _u64, _i64, _u32 = C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_int64, C.c_uint32


def oracle():
    global _o
    if _o is None:
        if not ORACLE_PATH.exists():
            raise ImportError("oracle/liboracle.so missing: run `make -C oracle`")
        lib = C.CDLL(str(ORACLE_PATH))
        sig = {
            "orc_apply_matrix": [_vp, _i, _i, _vp], "orc_cnot": [_vp, _i, _i, _i], "orc_pauli_x": [_vp, _i, _i],
            "orc_hadamard": [_vp, _i, _i], "orc_phase": [_vp, _i, _i, _d], "orc_pauli_y": [_vp, _i, _i],
            "orc_pauli_z": [_vp, _i, _i], "orc_rx": [_vp, _i, _i, _d], "orc_ry": [_vp, _i, _i, _d],
            "orc_rz": [_vp, _i, _i, _d], "orc_swap": [_vp, _i, _i, _i], "orc_mcphase": [_vp, _i, _u64, _i, _d],
            "orc_cphase": [_vp, _i, _i, _i, _d], "orc_sample": [_vp, _i, _vp, _i64, _vp],
            "orc_pauli_frame": [_vp, _i, _u64, _u64],
        }
        for k, a in sig.items():
            getattr(lib, k).argtypes = a
            getattr(lib, k).restype = None
        lib.orc_run_ops.argtypes = [_vp, _i, _vp, _i, _vp]
        lib.orc_run_ops.restype = _i
        lib.orc_measure.argtypes = [_vp, _i, _d]
        lib.orc_measure.restype = _u64
        lib.orc_collapse_qubit.argtypes = [_vp, _i, _i, _d, _i]
        lib.orc_collapse_qubit.restype = _i
        lib.orc_norm2.argtypes = [_vp, _i]
        lib.orc_norm2.restype = _d
        lib.orc_expect_zstrings.argtypes = [_vp, _i, _vp, _vp, _i]
        lib.orc_expect_zstrings.restype = _d
        lib.orc_run_measured.argtypes = [_vp, _i, _vp, _i, _vp, _vp, _i, _vp, _i]
        lib.orc_run_measured.restype = _i
        lib.orc_adjoint_expect.argtypes = [_i, _vp, _i, _vp, _i, _vp, _vp, _i, _vp, _vp]
        lib.orc_adjoint_expect.restype = _i
        lib.orc_noise_create.argtypes = [_i, _u32, _d]
        lib.orc_noise_create.restype = _vp
        lib.orc_noise_destroy.argtypes = [_vp]
        lib.orc_noise_set_rate.argtypes = [_vp, _i, _d]
        lib.orc_noise_accumulate.argtypes = [_vp, _i, _i]
        lib.orc_noise_masks.argtypes = [_vp, _vp, _vp, _vp]
        lib.orc_noise_reset.argtypes = [_vp]
        lib.orc_noise_trajectories.argtypes = [_i, _vp, _i, _u32, _i, _d, _d, _vp, _vp]
        lib.orc_mt19937_kat.restype = _u32
        lib.orc_uniforms.argtypes = [_u32, _i64, _vp]
        _o = lib
    return _o


def ref():
    """The reference's compiled CPU backends, or None when not built here."""
    global _r
    if _r is None and REF_PATH.exists():
        lib = C.CDLL(str(REF_PATH))
        lib.ref_run_ops.argtypes = [_i, _vp, _i, _vp, _i, _vp]
        lib.ref_run_ops.restype = _i
        lib.ref_time_ops.argtypes = [_i, _i, _vp, _vp, _i, _vp]
        lib.ref_time_ops.restype = _d
        lib.ref_noise_frame.argtypes = [_i, _vp, _i, _u32, _d, _vp, _vp, _vp, _vp, _vp]
        lib.ref_noise_frame.restype = _i
        _r = lib
    return _r


def _p(a):
    return None if a is None else a.ctypes.data


def run_ops(psi: np.ndarray, n: int, ops: np.ndarray, params=None) -> np.ndarray:
    s = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    pp = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    rc = oracle().orc_run_ops(_p(s), n, _p(ops), len(ops), _p(pp))
    assert rc == 0
    return s


def run_measured(psi: np.ndarray, n: int, ops: np.ndarray, u, nresults: int, params=None):
    """Gate list with MEASURE (13) / RESET_QUBIT (14) ops replayed with the
    uniforms u in list order: (state, results)."""
    s = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    pp = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    res = np.full(max(nresults, 1), -1, dtype=np.int32)
    rc = oracle().orc_run_measured(_p(s), n, _p(ops), len(ops), _p(pp), _p(uu), len(uu), _p(res), nresults)
    assert rc == 0
    return s, res[:nresults]


def ref_run_ops(psi: np.ndarray, n: int, ops: np.ndarray, params=None, backend: int = 0) -> np.ndarray:
    s = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    pp = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    rc = ref().ref_run_ops(backend, _p(s), n, _p(ops), len(ops), _p(pp))
    assert rc == 0
    return s


def basis(n: int, idx: int = 0) -> np.ndarray:
    s = np.zeros(1 << n, dtype=np.complex128)
    s[idx] = 1.0
    return s


def sample(psi: np.ndarray, n: int, u: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(psi, dtype=np.complex128)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty(len(uu), dtype=np.uint64)
    oracle().orc_sample(_p(s), n, _p(uu), len(uu), _p(out))
    return out


def measure(psi: np.ndarray, n: int, u: float):
    s = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    idx = oracle().orc_measure(_p(s), n, u)
    return int(idx), s


def collapse_qubit(psi, n, q, u, reset=False):
    s = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    o = oracle().orc_collapse_qubit(_p(s), n, q, u, int(reset))
    return int(o), s


def norm2(psi, n) -> float:
    return oracle().orc_norm2(_p(np.ascontiguousarray(psi, dtype=np.complex128)), n)


def expect_zstrings(psi, n, masks, weights) -> float:
    m = np.ascontiguousarray(masks, dtype=np.uint64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    return oracle().orc_expect_zstrings(_p(np.ascontiguousarray(psi, dtype=np.complex128)), n, _p(m), _p(w), len(m))


def adjoint_expect(n, ops, params, masks, weights):
    """Energy and adjoint gradient of sum_t w_t Z(m_t) after ops on |0...0>
    (plain sequential restatement, orc_adjoint_expect)."""
    m = np.ascontiguousarray(masks, dtype=np.uint64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    p = np.ascontiguousarray(params, dtype=np.float64)
    g = np.zeros(len(p))
    e = C.c_double(0.0)
    rc = oracle().orc_adjoint_expect(n, _p(ops), len(ops), _p(p), len(p), _p(m), _p(w), len(m), C.addressof(e), _p(g))
    assert rc == 0
    return e.value, g


def pauli_frame(psi, n, x, z):
    s = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    oracle().orc_pauli_frame(_p(s), n, x, z)
    return s


def noise_trajectories(n, ops, seed0, ntraj, rate=-1.0, mult=1.0):
    x = np.zeros(ntraj, dtype=np.uint64)
    z = np.zeros(ntraj, dtype=np.uint64)
    oracle().orc_noise_trajectories(n, _p(ops), len(ops), seed0, ntraj, rate, mult, _p(x), _p(z))
    return x, z


def uniforms(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    oracle().orc_uniforms(seed, count, _p(out))
    return out


def ref_noise_frame(n, ops, seed=42, mult=1.0):
    x = C.c_uint64(0)
    z = C.c_uint64(0)
    order = np.zeros(2 * n + 2, dtype=np.int32)
    norder = C.c_int(0)
    totals = np.zeros(2, dtype=np.int64)
    ref().ref_noise_frame(n, _p(ops), len(ops), seed, mult, C.addressof(x), C.addressof(z), _p(order),
                          C.addressof(norder), _p(totals))
    return x.value, z.value, order[: norder.value].tolist(), totals.tolist()
