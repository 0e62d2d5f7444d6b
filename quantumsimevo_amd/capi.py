"""ctypes binding of the C ABI in include/qse_hip.h (libqse_hip.so).

This is the Python-side view of the drop-in boundary: plain pointers and sizes,
status codes turned into `QSEError`. The library is loaded from the package
directory (built in-tree by `quantumsimevo_amd._build`); there is no fallback:
if it is missing, `load()` raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libqse_hip.so"

C128, C64 = 0, 1

# GateType numbering (reference Structurs.h:10) plus the generic 2x2 kind.
IDENTITY, X, Y, Z, H, PHASE, RX, RY, RZ, CNOT, CPHASE, MCPHASE, SWAP, MEASURE, RESET_QUBIT = range(15)
MATRIX = 32
KIND_NAMES = {IDENTITY: "I", X: "X", Y: "Y", Z: "Z", H: "H", PHASE: "PHASE", RX: "RX", RY: "RY", RZ: "RZ",
              CNOT: "CNOT", CPHASE: "CPHASE", MCPHASE: "MCPHASE", SWAP: "SWAP", MATRIX: "MATRIX",
              MEASURE: "MEASURE", RESET_QUBIT: "RESET"}

# struct qse_op (104 bytes)
OP_DTYPE = np.dtype([("kind", "<i4"), ("target", "<i4"), ("control", "<i4"), ("param", "<i4"),
                     ("controls", "<u8"), ("theta", "<f8"), ("scale", "<f8"), ("m", "<f8", (8,))])
assert OP_DTYPE.itemsize == 104

FAM_SWEEP, FAM_FUSED, FAM_READOUT, FAM_ADJOINT, FAM_BATCH, FAM_DENSE = range(6)
JIT_OFF, JIT_AUTO, JIT_ALWAYS = range(3)
ADJ_TERMS_AUTO, ADJ_TERMS_ROTATE, ADJ_TERMS_PAIRS, ADJ_TERMS_JOINT = range(4)

# Every symbol include/qse_hip.h declares (checked by the CPU tests).
EXPORTS = [
    "qse_abi_version", "qse_status_string", "qse_last_error", "qse_device_count", "qse_device_mem_info",
    "qse_device_info", "qse_ctx_create", "qse_ctx_destroy", "qse_ctx_sync", "qse_ctx_set_fusion",
    "qse_ctx_get_fusion", "qse_autotune", "qse_ctx_set_jit", "qse_ctx_jit_stats", "qse_ctx_jit_wait", "qse_ctx_enable_timing", "qse_ctx_timing", "qse_ctx_reset_timing",
    "qse_state_create", "qse_state_destroy", "qse_state_info", "qse_state_upload", "qse_state_download",
    "qse_state_download_range", "qse_state_set_basis", "qse_state_copy", "qse_state_set_random", "qse_apply_matrix",
    "qse_apply_hadamard", "qse_apply_x", "qse_apply_y", "qse_apply_z", "qse_apply_phase", "qse_apply_rx",
    "qse_apply_ry", "qse_apply_rz", "qse_apply_cnot", "qse_apply_cphase", "qse_apply_mcphase",
    "qse_apply_swap", "qse_pauli_frame", "qse_run", "qse_run_repeat", "qse_run_measured", "qse_run_sweeps", "qse_plan_passes", "qse_plan_info", "qse_qft_probe", "qse_copy_bandwidth", "qse_exchange_probe", "qse_sample",
    "qse_measure", "qse_norm2", "qse_prob_one", "qse_collapse_qubit", "qse_expect_zstrings", "qse_expect_zstrings_noalloc",
    "qse_adjoint_expect", "qse_batch_expect", "qse_batch_run_small", "qse_sample_frames",
    "qse_compose_dense", "qse_apply_dense", "qse_state_renormalize", "qse_ctx_set_renormalize",
    "qse_ctx_set_adjoint_terms", "qse_batch_autotune", "qse_ctx_set_dense_blocks",
]


class QSEError(RuntimeError):
    """A non-zero status from libqse_hip (message from qse_last_error)."""

    def __init__(self, status: int, func: str, msg: str):
        super().__init__(f"{func}: status {status}: {msg}")
        self.status = status


_lib = None

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_u64 = C.c_uint64
_i64 = C.c_int64
_pd = C.POINTER(C.c_double)
_pu64 = C.POINTER(C.c_uint64)


def _sigs(lib):
    S = {
        "qse_abi_version": ([], _i),
        "qse_status_string": ([_i], C.c_char_p),
        "qse_last_error": ([], C.c_char_p),
        "qse_device_count": ([C.POINTER(_i)], _i),
        "qse_device_mem_info": ([_i, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], _i),
        "qse_device_info": ([_i, C.c_char_p, C.c_size_t, C.POINTER(_i), C.POINTER(C.c_size_t)], _i),
        "qse_ctx_create": ([_i, C.POINTER(_vp)], _i),
        "qse_ctx_destroy": ([_vp], _i),
        "qse_ctx_sync": ([_vp], _i),
        "qse_ctx_set_fusion": ([_vp, _i, _i, _i], _i),
        "qse_ctx_get_fusion": ([_vp, _i, C.POINTER(_i), C.POINTER(_i)], _i),
        "qse_ctx_set_jit": ([_vp, _i], _i),
        "qse_ctx_jit_stats": ([_vp, C.POINTER(_i64), C.POINTER(_i64), _pd], _i),
        "qse_ctx_jit_wait": ([_vp], _i),
        "qse_autotune": ([_vp, _vp, _vp, _i, _vp, _i, C.POINTER(_i), C.POINTER(_i), _pd, _vp, _vp, _vp,
                          C.POINTER(_i)], _i),
        "qse_ctx_enable_timing": ([_vp, _i], _i),
        "qse_ctx_timing": ([_vp, _i, C.POINTER(_i64), _pd, _pd], _i),
        "qse_ctx_reset_timing": ([_vp], _i),
        "qse_state_create": ([_vp, _i, _i, C.POINTER(_vp)], _i),
        "qse_state_destroy": ([_vp, _vp], _i),
        "qse_state_info": ([_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_vp), C.POINTER(C.c_size_t)], _i),
        "qse_state_upload": ([_vp, _vp, _vp], _i),
        "qse_state_download": ([_vp, _vp, _vp], _i),
        "qse_state_download_range": ([_vp, _vp, _u64, _u64, _vp], _i),
        "qse_state_set_basis": ([_vp, _vp, _u64], _i),
        "qse_state_copy": ([_vp, _vp, _vp], _i),
        "qse_state_set_random": ([_vp, _vp, _u64], _i),
        "qse_apply_matrix": ([_vp, _vp, _i, _pd], _i),
        "qse_apply_hadamard": ([_vp, _vp, _i], _i),
        "qse_apply_x": ([_vp, _vp, _i], _i),
        "qse_apply_y": ([_vp, _vp, _i], _i),
        "qse_apply_z": ([_vp, _vp, _i], _i),
        "qse_apply_phase": ([_vp, _vp, _i, _d], _i),
        "qse_apply_rx": ([_vp, _vp, _i, _d], _i),
        "qse_apply_ry": ([_vp, _vp, _i, _d], _i),
        "qse_apply_rz": ([_vp, _vp, _i, _d], _i),
        "qse_apply_cnot": ([_vp, _vp, _i, _i], _i),
        "qse_apply_cphase": ([_vp, _vp, _i, _i, _d], _i),
        "qse_apply_mcphase": ([_vp, _vp, C.POINTER(_i), _i, _i, _d], _i),
        "qse_apply_swap": ([_vp, _vp, _i, _i], _i),
        "qse_pauli_frame": ([_vp, _vp, _u64, _u64], _i),
        "qse_run": ([_vp, _vp, _vp, _i, _vp], _i),
        "qse_run_sweeps": ([_vp, _vp, _vp, _i, _vp], _i),
        "qse_run_repeat": ([_vp, _vp, _vp, _i, _vp, _i], _i),
        "qse_run_measured": ([_vp, _vp, _vp, _i, _vp, _vp, _i, _vp, _i], _i),
        "qse_plan_passes": ([_vp, _i, _i, _vp, _i, C.POINTER(_i)], _i),
        "qse_plan_info": ([_vp, _i, _i, _vp, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)], _i),
        "qse_qft_probe": ([_vp, _i, C.c_uint64, _vp, _i, _vp], _i),
        "qse_copy_bandwidth": ([_vp, _i, _i, _pd], _i),
        "qse_exchange_probe": ([_vp, _vp, _i, _i, _pd], _i),
        "qse_compose_dense": ([_i, _vp, _i, _vp, _vp], _i),
        "qse_state_renormalize": ([_vp, _vp, _vp], _i),
        "qse_ctx_set_renormalize": ([_vp, _i], _i),
        "qse_ctx_set_adjoint_terms": ([_vp, _i], _i),
        "qse_batch_autotune": ([_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, C.POINTER(_i), C.POINTER(_i),
                                C.POINTER(_i), _pd], _i),
        "qse_apply_dense": ([_vp, _vp, _i, _vp], _i),
        "qse_ctx_set_dense_blocks": ([_vp, _i], _i),
        "qse_sample": ([_vp, _vp, _vp, _i64, _vp], _i),
        "qse_measure": ([_vp, _vp, _d, _pu64], _i),
        "qse_norm2": ([_vp, _vp, _pd], _i),
        "qse_prob_one": ([_vp, _vp, _i, _pd], _i),
        "qse_collapse_qubit": ([_vp, _vp, _i, _d, _i, C.POINTER(_i)], _i),
        "qse_expect_zstrings": ([_vp, _vp, _vp, _vp, _i, _pd], _i),
        "qse_expect_zstrings_noalloc": ([_vp, _vp, _vp, _vp, _i, _pd], _i),
        "qse_adjoint_expect": ([_vp, _i, _i, _vp, _i, _vp, _i, _i, _vp, _vp, _i, _i, _vp, _vp], _i),
        "qse_batch_expect": ([_vp, _i, _i, _vp, _i, _vp, _i, _i, _vp, _vp, _i, _vp], _i),
        "qse_batch_run_small": ([_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp], _i),
        "qse_sample_frames": ([_vp, _vp, _vp, _vp, _i, _vp, _i64, _vp], _i),
    }
    for name, (args, res) in S.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res


def load():
    """Load libqse_hip.so from the package directory (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built: run quantumsimevo_amd._build.build_all()")
        lib = C.CDLL(str(LIB_PATH))
        _sigs(lib)
        _lib = lib
    return _lib


def check(status: int, func: str = "qse") -> None:
    if status != 0:
        lib = load()
        raise QSEError(status, func, lib.qse_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def device_count() -> int:
    c = C.c_int(0)
    check(load().qse_device_count(C.byref(c)), "qse_device_count")
    return c.value


def device_info(device: int = 0) -> dict:
    name = C.create_string_buffer(256)
    cus = C.c_int(0)
    tot = C.c_size_t(0)
    check(load().qse_device_info(device, name, 256, C.byref(cus), C.byref(tot)), "qse_device_info")
    free = C.c_size_t(0)
    tot2 = C.c_size_t(0)
    check(load().qse_device_mem_info(device, C.byref(free), C.byref(tot2)), "qse_device_mem_info")
    return {"name": name.value.decode(), "compute_units": cus.value, "total_bytes": tot.value,
            "free_bytes": free.value}


def compose_dense(m: int, ops, params=None) -> np.ndarray:
    """The (2^m, 2^m) unitary of a gate list on qubits 0..m-1 (host computation)."""
    a = as_ops(ops)
    p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    u = np.empty((1 << m, 1 << m), dtype=np.complex128)
    check(load().qse_compose_dense(int(m), _ptr(a), len(a), None if p is None else _ptr(p), _ptr(u)),
          "qse_compose_dense")
    return u


def as_ops(ops) -> np.ndarray:
    a = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    return a


class Context:
    """One device, one HIP stream (qse_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.qse_ctx_create(device, C.byref(h)), "qse_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.qse_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.lib.qse_ctx_sync(self.h), "qse_ctx_sync")

    def set_fusion(self, precision: int, tile_bits: int, low_bits: int):
        check(self.lib.qse_ctx_set_fusion(self.h, precision, tile_bits, low_bits), "qse_ctx_set_fusion")

    def get_fusion(self, precision: int):
        k, l = C.c_int(0), C.c_int(0)
        check(self.lib.qse_ctx_get_fusion(self.h, precision, C.byref(k), C.byref(l)), "qse_ctx_get_fusion")
        return k.value, l.value

    def set_adjoint_terms(self, mode: int):
        """ADJ_TERMS_AUTO (default), ADJ_TERMS_ROTATE, ADJ_TERMS_PAIRS or
        ADJ_TERMS_JOINT: how the adjoint evaluates RX/RY generator terms on 13+
        qubit states (JOINT: the reverse sweep as compiled joint passes over psi
        and lambda with every generator term summed in the tile)."""
        check(self.lib.qse_ctx_set_adjoint_terms(self.h, int(mode)), "qse_ctx_set_adjoint_terms")

    def set_dense_blocks(self, ops_per_qubit: int):
        """qse_run's routing of deep blocks on qubits 0..5 to the matrix cores
        (threshold in ops per qubit of the block; 0 = never)."""
        check(self.lib.qse_ctx_set_dense_blocks(self.h, ops_per_qubit), "qse_ctx_set_dense_blocks")

    def set_renormalize(self, every_passes: int):
        """Renormalise states (FP64, on the device) once qse_run calls have made
        every_passes passes over them since the last time; 0 turns it off."""
        check(self.lib.qse_ctx_set_renormalize(self.h, int(every_passes)), "qse_ctx_set_renormalize")

    def set_jit(self, mode: int):
        """JIT_OFF, JIT_AUTO (default: compile a program on its second run) or JIT_ALWAYS."""
        check(self.lib.qse_ctx_set_jit(self.h, mode), "qse_ctx_set_jit")

    def jit_stats(self) -> dict:
        c, l, ms = C.c_int64(0), C.c_int64(0), C.c_double(0)
        check(self.lib.qse_ctx_jit_stats(self.h, C.byref(c), C.byref(l), C.byref(ms)), "qse_ctx_jit_stats")
        return {"compiled": c.value, "launches": l.value, "compile_ms": ms.value}

    def autotune(self, state: "State", ops, params=None, candidates: int = 3) -> dict:
        """Time the tuner's best `candidates` tile shapes on `state` (overwritten) and keep the fastest
        for this gate list's structure; returns {"k", "L", "ms", "tried": [(k, L, ms), ...]}."""
        a = as_ops(ops)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        k, L, ms, nt = C.c_int(0), C.c_int(0), C.c_double(0), C.c_int(0)
        tk = np.zeros(candidates, np.int32)
        tl = np.zeros(candidates, np.int32)
        tm = np.zeros(candidates, np.float64)
        check(self.lib.qse_autotune(self.h, state.h, _ptr(a), len(a), None if p is None else _ptr(p), candidates,
                                    C.byref(k), C.byref(L), C.byref(ms), _ptr(tk), _ptr(tl), _ptr(tm), C.byref(nt)),
              "qse_autotune")
        return {"k": k.value, "L": L.value, "ms": ms.value,
                "tried": [(int(tk[i]), int(tl[i]), float(tm[i])) for i in range(nt.value)]}

    def jit_wait(self):
        """Wait for the structure compilations running on worker threads."""
        check(self.lib.qse_ctx_jit_wait(self.h), "qse_ctx_jit_wait")

    def enable_timing(self, on: bool = True):
        check(self.lib.qse_ctx_enable_timing(self.h, int(on)), "qse_ctx_enable_timing")

    def reset_timing(self):
        check(self.lib.qse_ctx_reset_timing(self.h), "qse_ctx_reset_timing")

    def timing(self, family: int):
        n = C.c_int64(0)
        ms = C.c_double(0)
        b = C.c_double(0)
        check(self.lib.qse_ctx_timing(self.h, family, C.byref(n), C.byref(ms), C.byref(b)), "qse_ctx_timing")
        return {"launches": n.value, "ms": ms.value, "bytes": b.value}

    def state(self, nqubits: int, precision: int = C128) -> "State":
        return State(self, nqubits, precision)

    def plan_passes(self, nqubits: int, precision: int, ops) -> int:
        a = as_ops(ops)
        p = C.c_int(0)
        check(self.lib.qse_plan_passes(self.h, nqubits, precision, _ptr(a), len(a), C.byref(p)), "qse_plan_passes")
        return p.value

    def plan_info(self, nqubits: int, precision: int, ops) -> dict:
        """The plan qse_run would execute: passes, tile shape and LDS stages."""
        a = as_ops(ops)
        v = [C.c_int(0) for _ in range(4)]
        check(self.lib.qse_plan_info(self.h, nqubits, precision, _ptr(a), len(a), *[C.byref(x) for x in v]),
              "qse_plan_info")
        return {"passes": v[0].value, "tile_bits": v[1].value, "low_bits": v[2].value, "stages": v[3].value}

    def qft_probe(self, nqubits: int, seed: int, indices) -> np.ndarray:
        """FP64 closed-form <k|QFT|psi_seed> for the normalised random state of
        State.set_random(seed) (complex128 array, one entry per index)."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        out = np.empty(2 * len(idx))
        check(self.lib.qse_qft_probe(self.h, nqubits, seed, _ptr(idx), len(idx), _ptr(out)), "qse_qft_probe")
        return out.view(np.complex128)

    def copy_bandwidth(self, log2_elems: int = 30, run_log2: int = 12) -> float:
        """GB/s of a tiled out-of-place copy with contiguous runs of 2^run_log2 16-byte elements."""
        g = C.c_double(0)
        check(self.lib.qse_copy_bandwidth(self.h, log2_elems, run_log2, C.byref(g)), "qse_copy_bandwidth")
        return g.value

    # ---- batched entry points ----
    def batch_run_small(self, precision, circuits, obs_masks, obs_weights, grads: bool = True):
        """Many circuits [(n, ops, params), ...] of uneven sizes in persistent
        launches (forward; with grads, the reverse sweeps): energies (ncirc,)
        and, with grads, a list of gradient vectors."""
        from .circuits import pack_batch
        nq, op_off, ops, par_off, params = pack_batch(circuits)
        ops = as_ops(ops) if len(ops) else np.zeros(0, dtype=OP_DTYPE)
        m = np.ascontiguousarray(obs_masks, dtype=np.uint64)
        w = np.ascontiguousarray(obs_weights, dtype=np.float64)
        E = np.empty(len(circuits))
        G = np.empty(max(len(params), 1)) if grads else None
        check(self.lib.qse_batch_run_small(self.h, precision, len(circuits), _ptr(nq), _ptr(op_off), _ptr(ops),
                                           _ptr(par_off), _ptr(params) if len(params) else None, _ptr(m), _ptr(w),
                                           len(m), _ptr(E), _ptr(G) if grads else None), "qse_batch_run_small")
        if not grads:
            return E
        return E, [G[par_off[i]:par_off[i + 1]].copy() for i in range(len(circuits))]
    def batch_autotune(self, precision, circuits, obs_masks, obs_weights, grads: bool = True) -> dict:
        """Time the batch engine's tile shapes on these circuits (a sample of a
        larger batch) and keep the fastest for later batch_run_small calls."""
        from .circuits import pack_batch
        nq, op_off, ops, par_off, params = pack_batch(circuits)
        ops = as_ops(ops) if len(ops) else np.zeros(0, dtype=OP_DTYPE)
        m = np.ascontiguousarray(obs_masks, dtype=np.uint64)
        w = np.ascontiguousarray(obs_weights, dtype=np.float64)
        k, L, T, ms = C.c_int(0), C.c_int(0), C.c_int(0), C.c_double(0)
        check(self.lib.qse_batch_autotune(self.h, precision, len(circuits), _ptr(nq), _ptr(op_off), _ptr(ops),
                                          _ptr(par_off), _ptr(params) if len(params) else None, _ptr(m), _ptr(w),
                                          len(m), int(grads), C.byref(k), C.byref(L), C.byref(T), C.byref(ms)),
              "qse_batch_autotune")
        return {"tile_qubits": k.value, "low_qubits": L.value, "team_workgroups": T.value, "ms": ms.value}

    def adjoint_expect(self, nqubits, precision, ops, params, obs_masks, obs_weights, ckpt_every=0):
        a = as_ops(ops)
        P = np.ascontiguousarray(params, dtype=np.float64)
        if P.ndim == 1:
            P = P[None, :]
        B, npar = P.shape
        m = np.ascontiguousarray(obs_masks, dtype=np.uint64)
        w = np.ascontiguousarray(obs_weights, dtype=np.float64)
        E = np.zeros(B)
        G = np.zeros((B, npar))
        check(self.lib.qse_adjoint_expect(self.h, nqubits, precision, _ptr(a), len(a), _ptr(P), npar, B, _ptr(m),
                                          _ptr(w), len(m), ckpt_every, _ptr(E), _ptr(G)), "qse_adjoint_expect")
        return E, G

    def batch_expect(self, nqubits, precision, ops, params, obs_masks, obs_weights):
        a = as_ops(ops)
        P = np.ascontiguousarray(params, dtype=np.float64)
        if P.ndim == 1:
            P = P[None, :]
        B, npar = P.shape
        m = np.ascontiguousarray(obs_masks, dtype=np.uint64)
        w = np.ascontiguousarray(obs_weights, dtype=np.float64)
        E = np.zeros(B)
        check(self.lib.qse_batch_expect(self.h, nqubits, precision, _ptr(a), len(a), _ptr(P), npar, B, _ptr(m),
                                        _ptr(w), len(m), _ptr(E)), "qse_batch_expect")
        return E


class State:
    """A device state vector (qse_state). Host views are complex128 numpy arrays."""

    def __init__(self, ctx: Context, nqubits: int, precision: int = C128):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n = nqubits
        self.precision = precision
        h = C.c_void_p()
        check(self.lib.qse_state_create(ctx.h, nqubits, precision, C.byref(h)), "qse_state_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.qse_state_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self) -> int:
        return 1 << self.n

    def device_ptr(self) -> int:
        p = C.c_void_p()
        check(self.lib.qse_state_info(self.h, None, None, C.byref(p), None), "qse_state_info")
        return p.value

    def upload(self, amps: np.ndarray):
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        if a.size != self.size:
            raise ValueError("size mismatch")
        check(self.lib.qse_state_upload(self.ctx.h, self.h, _ptr(a)), "qse_state_upload")

    def download(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.complex128)
        check(self.lib.qse_state_download(self.ctx.h, self.h, _ptr(out)), "qse_state_download")
        return out

    def download_range(self, first: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.complex128)
        check(self.lib.qse_state_download_range(self.ctx.h, self.h, first, count, _ptr(out)),
              "qse_state_download_range")
        return out

    def set_basis(self, index: int = 0):
        check(self.lib.qse_state_set_basis(self.ctx.h, self.h, index), "qse_state_set_basis")

    def set_random(self, seed: int):
        check(self.lib.qse_state_set_random(self.ctx.h, self.h, seed), "qse_state_set_random")

    def copy_from(self, other: "State"):
        check(self.lib.qse_state_copy(self.ctx.h, self.h, other.h), "qse_state_copy")

    # single gates
    def matrix(self, q, m):
        mm = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(4).view(np.float64))
        check(self.lib.qse_apply_matrix(self.ctx.h, self.h, q, mm.ctypes.data_as(_pd)), "qse_apply_matrix")

    def h_(self, q): check(self.lib.qse_apply_hadamard(self.ctx.h, self.h, q), "qse_apply_hadamard")
    def x(self, q): check(self.lib.qse_apply_x(self.ctx.h, self.h, q), "qse_apply_x")
    def y(self, q): check(self.lib.qse_apply_y(self.ctx.h, self.h, q), "qse_apply_y")
    def z(self, q): check(self.lib.qse_apply_z(self.ctx.h, self.h, q), "qse_apply_z")
    def phase(self, q, t): check(self.lib.qse_apply_phase(self.ctx.h, self.h, q, t), "qse_apply_phase")
    def rx(self, q, t): check(self.lib.qse_apply_rx(self.ctx.h, self.h, q, t), "qse_apply_rx")
    def ry(self, q, t): check(self.lib.qse_apply_ry(self.ctx.h, self.h, q, t), "qse_apply_ry")
    def rz(self, q, t): check(self.lib.qse_apply_rz(self.ctx.h, self.h, q, t), "qse_apply_rz")
    def cnot(self, c, t): check(self.lib.qse_apply_cnot(self.ctx.h, self.h, c, t), "qse_apply_cnot")
    def cphase(self, c, t, th): check(self.lib.qse_apply_cphase(self.ctx.h, self.h, c, t, th), "qse_apply_cphase")

    def mcphase(self, controls, t, th):
        cs = (C.c_int * len(controls))(*controls)
        check(self.lib.qse_apply_mcphase(self.ctx.h, self.h, cs, len(controls), t, th), "qse_apply_mcphase")

    def swap(self, a, b): check(self.lib.qse_apply_swap(self.ctx.h, self.h, a, b), "qse_apply_swap")

    def pauli_frame(self, xmask, zmask):
        check(self.lib.qse_pauli_frame(self.ctx.h, self.h, xmask, zmask), "qse_pauli_frame")

    def run(self, ops, params=None, fused: bool = True):
        a = as_ops(ops)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        f = self.lib.qse_run if fused else self.lib.qse_run_sweeps
        check(f(self.ctx.h, self.h, _ptr(a), len(a), None if p is None else _ptr(p)),
              "qse_run" if fused else "qse_run_sweeps")

    def run_measured(self, ops, u, nresults: int, params=None) -> np.ndarray:
        """A gate list with MEASURE / RESET_QUBIT ops (qse_run_measured): one
        uniform of u per measurement op in list order; returns the outcomes by
        result slot (-1 for slots no MEASURE op wrote)."""
        a = as_ops(ops)
        uu = np.ascontiguousarray(u, dtype=np.float64)
        res = np.full(max(nresults, 1), -1, dtype=np.int32)
        pp = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        check(self.lib.qse_run_measured(self.ctx.h, self.h, _ptr(a), len(a), _ptr(pp) if pp is not None else None,
                                        _ptr(uu) if len(uu) else None, len(uu), _ptr(res), nresults),
              "qse_run_measured")
        return res[:nresults]

    def run_repeat(self, ops, reps: int, params=None):
        """The gate list applied reps times (replayed from a HIP graph after two eager runs)."""
        a = as_ops(ops)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        check(self.lib.qse_run_repeat(self.ctx.h, self.h, _ptr(a), len(a), None if p is None else _ptr(p), int(reps)),
              "qse_run_repeat")

    def renormalize(self) -> float:
        """psi / ||psi|| in FP64 on the device; returns the norm^2 before."""
        v = C.c_double(0)
        check(self.lib.qse_state_renormalize(self.ctx.h, self.h, C.byref(v)), "qse_state_renormalize")
        return v.value

    def apply_dense(self, m: int, u):
        """psi <- (U on qubits 0..m-1) psi with the FP64 matrix cores (m = 4, 5, 6;
        complex128 states); u is a complex (2^m, 2^m) matrix."""
        uu = np.ascontiguousarray(u, dtype=np.complex128)
        if uu.shape != (1 << m, 1 << m):
            raise ValueError("u must be a (2^m, 2^m) matrix")
        check(self.lib.qse_apply_dense(self.ctx.h, self.h, int(m), _ptr(uu)), "qse_apply_dense")

    # readout
    def exchange_probe(self, method: int, reps: int) -> float:
        """reps layers of RY(0.3 + 0.1 q) on qubits 0..5 applied in place in one pass,
        pairs exchanged through LDS (method 0) or across wavefront lanes by
        DPP/ds_swizzle (method 1); returns the launch's milliseconds."""
        ms = C.c_double(0)
        check(self.lib.qse_exchange_probe(self.ctx.h, self.h, method, reps, C.byref(ms)), "qse_exchange_probe")
        return ms.value

    def sample(self, u) -> np.ndarray:
        uu = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty(len(uu), dtype=np.uint64)
        check(self.lib.qse_sample(self.ctx.h, self.h, _ptr(uu), len(uu), _ptr(out)), "qse_sample")
        return out

    def sample_frames(self, xmasks, u) -> np.ndarray:
        xm = np.ascontiguousarray(xmasks, dtype=np.uint64)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(len(xm), -1)
        out = np.empty(uu.shape, dtype=np.uint64)
        check(self.lib.qse_sample_frames(self.ctx.h, self.h, _ptr(xm), None, len(xm), _ptr(uu), uu.shape[1],
                                         _ptr(out)), "qse_sample_frames")
        return out

    def measure(self, u: float) -> int:
        o = C.c_uint64(0)
        check(self.lib.qse_measure(self.ctx.h, self.h, u, C.byref(o)), "qse_measure")
        return o.value

    def norm2(self) -> float:
        v = C.c_double(0)
        check(self.lib.qse_norm2(self.ctx.h, self.h, C.byref(v)), "qse_norm2")
        return v.value

    def prob_one(self, q) -> float:
        v = C.c_double(0)
        check(self.lib.qse_prob_one(self.ctx.h, self.h, q, C.byref(v)), "qse_prob_one")
        return v.value

    def collapse_qubit(self, q, u, reset=False) -> int:
        o = C.c_int(0)
        check(self.lib.qse_collapse_qubit(self.ctx.h, self.h, q, u, int(reset), C.byref(o)), "qse_collapse_qubit")
        return o.value

    def expect_zstrings(self, masks, weights) -> float:
        m = np.ascontiguousarray(masks, dtype=np.uint64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        v = C.c_double(0)
        check(self.lib.qse_expect_zstrings(self.ctx.h, self.h, _ptr(m), _ptr(w), len(m), C.byref(v)),
              "qse_expect_zstrings")
        return v.value

    def expect_zstrings_noalloc(self, masks, weights) -> float:
        """expect_zstrings through the allocation-free entry point (at most 64
        terms; a fixed per-device scratch array instead of allocated buffers)."""
        m = np.ascontiguousarray(masks, dtype=np.uint64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        v = C.c_double(0)
        check(self.lib.qse_expect_zstrings_noalloc(self.ctx.h, self.h, _ptr(m), _ptr(w), len(m), C.byref(v)),
              "qse_expect_zstrings_noalloc")
        return v.value
