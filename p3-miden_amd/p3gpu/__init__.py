"""Python host mirror of the reference's lifted-PCS interfaces, over the C ABI
of libp3gpu.so (include/p3gpu.h).

The names follow the reference crates so that callers and tests read like the
reference's own code:

* ``Dft.coset_lde_batch``            — p3-dft TwoAdicSubgroupDft (commit.rs:217-221)
* ``LmcsConfig.build_tree`` / ``build_aligned_tree`` / ``hash`` / ``compress``
                                     — p3-miden-lmcs/src/lmcs.rs:102-149
* ``LiftedMerkleTree.root/height/rows/widths/prove_batch``
                                     — p3-miden-lmcs/src/lifted_tree.rs:102-224
* ``ProverTranscript``               — p3-miden-transcript/src/prover.rs:15-145
* ``PcsParams``                      — p3-miden-lifted-fri/src/params.rs:170-204
* ``open_with_channel``              — p3-miden-lifted-fri/src/prover.rs:35-116
* ``commit_traces``                  — p3-miden-lifted-stark/src/prover/commit.rs:175-228

Every computation runs in the HIP library; there is no CPU fallback.  If the
library is missing or no GPU is visible, the constructors raise.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_PKG, "libp3gpu.so")

P = 0xFFFFFFFF00000001
GENERATOR = 7
HASH_POSEIDON2 = 0
HASH_KECCAK = 1
CHALLENGER_DUPLEX_P2 = 0
CHALLENGER_KECCAK = 1
DEVICE_PTRS = 1
BITREV_ROWS = 2
TAKE_OWNERSHIP = 4

_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p


class P3GError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        detail = _lib.p3g_last_error().decode() if _lib is not None else ""
        msg = _lib.p3g_status_string(code).decode() if _lib is not None else str(code)
        super().__init__(f"{where}: {msg} ({code}){': ' + detail if detail else ''}")


class _Matrix(C.Structure):
    _fields_ = [("data", _u64p), ("height", C.c_uint64), ("width", C.c_uint64)]


class PcsParams(C.Structure):
    """PcsParams::new(log_blowup, log_folding_arity, log_final_degree, folding_pow_bits,
    deep_pow_bits, num_queries, query_pow_bits) with the reference's validation
    (p3-miden-lifted-fri/src/params.rs:170-204)."""

    _fields_ = [(n, C.c_uint32) for n in ("log_blowup", "log_folding_arity", "log_final_degree",
                                           "folding_pow_bits", "deep_pow_bits", "num_queries", "query_pow_bits")]

    def __init__(self, log_blowup, log_folding_arity, log_final_degree, folding_pow_bits=0, deep_pow_bits=0,
                 num_queries=1, query_pow_bits=0):
        if log_folding_arity not in (1, 2, 3):
            raise ValueError(f"invalid folding arity: log_arity {log_folding_arity} (must be 1, 2, or 3)")
        if log_blowup == 0:
            raise ValueError("log_blowup must be > 0")
        if log_final_degree + log_blowup > 64:
            raise ValueError("log_final_degree + log_blowup exceeds 64")
        if num_queries == 0:
            raise ValueError("num_queries must be > 0")
        if max(folding_pow_bits, deep_pow_bits, query_pow_bits) >= 32:
            raise ValueError("proof-of-work bits must be below 32 (the library's grinding bound)")
        super().__init__(log_blowup, log_folding_arity, log_final_degree, folding_pow_bits, deep_pow_bits,
                         num_queries, query_pow_bits)

    def num_rounds(self, log_domain: int) -> int:
        lmf = self.log_final_degree + self.log_blowup
        diff = max(0, log_domain - lmf)
        return -(-diff // self.log_folding_arity)

    def final_poly_degree(self, log_domain: int) -> int:
        lfs = log_domain - self.num_rounds(log_domain) * self.log_folding_arity
        return 1 << max(0, lfs - self.log_blowup)


_SEND = C.CFUNCTYPE(C.c_int, C.c_void_p, _u64p, C.c_uint64)
_COMM = C.CFUNCTYPE(C.c_int, C.c_void_p, _u64p)
_GRIND = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, _u64p)
_SAMPLE = C.CFUNCTYPE(C.c_int, C.c_void_p, _u64p)
_SBITS = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, _u64p)


class _Channel(C.Structure):
    _fields_ = [("user", C.c_void_p), ("send_fields", _SEND), ("send_commitment", _COMM), ("hint_fields", _SEND),
                ("hint_commitment", _COMM), ("grind", _GRIND), ("sample", _SAMPLE), ("sample_bits", _SBITS)]


class CallbackChannel:
    """Adapts a caller-side prover channel (any object with send_field_slice,
    send_commitment, hint_field_slice, hint_commitment, grind, sample and
    sample_bits) to the library's p3g_channel callbacks."""

    def __init__(self, channel):
        self.channel = channel
        ch = channel

        def arr(ptr, n):
            return [int(ptr[i]) for i in range(n)]

        def wrap(fn):
            def inner(*args):
                try:
                    return fn(*args) or 0
                except Exception:
                    return 1
            return inner

        self._fns = (
            _SEND(wrap(lambda u, p, n: ch.send_field_slice(arr(p, n)))),
            _COMM(wrap(lambda u, p: ch.send_commitment(arr(p, 4)))),
            _SEND(wrap(lambda u, p, n: ch.hint_field_slice(arr(p, n)))),
            _COMM(wrap(lambda u, p: ch.hint_commitment(arr(p, 4)))),
            _GRIND(wrap(lambda u, b, out: out.__setitem__(0, ch.grind(b)))),
            _SAMPLE(wrap(lambda u, out: out.__setitem__(0, ch.sample()))),
            _SBITS(wrap(lambda u, b, out: out.__setitem__(0, ch.sample_bits(b)))),
        )
        self.struct = _Channel(None, *self._fns)


class _Request(C.Structure):
    _fields_ = [("hash_kind", C.c_uint32), ("params", PcsParams), ("traces", C.POINTER(_Matrix)),
                ("k", C.c_uint64), ("flags", C.c_uint32)]


class _ResultInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("root", C.c_uint64 * 4), ("eval_points", C.c_uint64 * 4),
                ("num_fields", C.c_uint64), ("num_commitments", C.c_uint64), ("submit_ms", C.c_double),
                ("admit_ms", C.c_double), ("done_ms", C.c_double), ("worker", C.c_uint32),
                ("footprint_bytes", C.c_uint64)]


class _ServerCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("admitted", "completed", "queued", "running", "batches", "batched_requests",
                                          "inflight_bytes",
                                          "peak_inflight_bytes", "peak_running", "budget_bytes", "n_malloc",
                                          "n_free")] + [("malloc_ms", C.c_double), ("free_ms", C.c_double)]


def _load():
    if not os.path.exists(LIB_PATH):
        return None
    lib = C.CDLL(LIB_PATH)
    sig = {
        "p3g_status_string": (C.c_char_p, [C.c_int]),
        "p3g_last_error": (C.c_char_p, []),
        "p3g_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
        "p3g_ctx_create_on_stream": (C.c_int, [C.c_int, _vp, C.POINTER(_vp)]),
        "p3g_ctx_destroy": (None, [_vp]),
        "p3g_ctx_synchronize": (C.c_int, [_vp]),
        "p3g_ctx_trim": (C.c_int, [_vp]),
        "p3g_ctx_memory": (C.c_int, [_vp, _u64p, _u64p, _u64p, C.c_int]),
        "p3g_profile_enable": (C.c_int, [_vp, C.c_int]),
        "p3g_profile_reset": (C.c_int, [_vp]),
        "p3g_profile_read": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_double), _u64p]),
        "p3g_calibrate_valu": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double), _u64p]),
        "p3g_field_check": (C.c_int, [_vp, C.c_uint32, C.c_uint32, _u64p, _u64p, _u64p, C.c_uint64, _u64p, _u64p]),
        "p3g_calibrate_copy": (C.c_int, [_vp, C.c_uint64, C.c_uint32, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]),
        "p3g_alloc": (C.c_int, [_vp, C.c_uint64, C.POINTER(_vp)]),
        "p3g_free": (C.c_int, [_vp, _vp]),
        "p3g_upload": (C.c_int, [_vp, _vp, _vp, C.c_uint64]),
        "p3g_download": (C.c_int, [_vp, _vp, _vp, C.c_uint64]),
        "p3g_copy_device": (C.c_int, [_vp, _vp, _vp, C.c_uint64]),
        "p3g_fill_random": (C.c_int, [C.c_uint64, C.c_uint64, _u64p]),
        "p3g_coset_lde_batch": (C.c_int, [_vp, _u64p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, _u64p]),
        "p3g_coset_lde_batch_noalloc": (C.c_int, [_vp, _vp, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint64, _vp]),
        "p3g_eval_constraints": (C.c_int, [_vp, _vp, C.c_uint64, C.POINTER(_Matrix), C.POINTER(_Matrix), _u64p,
                                           C.c_uint64, C.c_uint64, _u64p, C.c_uint64, _u64p, C.c_uint64,
                                           C.c_uint64, C.c_uint32, C.c_uint64, _u64p, _vp]),
        "p3g_air_compile": (C.c_int, [_vp, C.c_uint64, _u64p]),
        "p3g_lmcs_alignment": (C.c_uint32, [C.c_uint32]),
        "p3g_lmcs_build": (C.c_int, [_vp, C.c_uint32, C.POINTER(_Matrix), C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
        "p3g_lmcs_hash": (C.c_int, [C.c_uint32, _u64p, _u64p, C.c_uint64, _u64p]),
        "p3g_lmcs_compress": (C.c_int, [C.c_uint32, _u64p, _u64p, _u64p]),
        "p3g_tree_root": (C.c_int, [_vp, _u64p]),
        "p3g_tree_height": (C.c_uint64, [_vp]),
        "p3g_tree_alignment": (C.c_uint32, [_vp]),
        "p3g_tree_num_matrices": (C.c_uint64, [_vp]),
        "p3g_tree_widths": (C.c_int, [_vp, _u64p]),
        "p3g_tree_rows": (C.c_int, [_vp, C.c_uint64, _u64p, C.c_uint64, _u64p]),
        "p3g_lmcs_build_salted": (C.c_int, [_vp, C.c_uint32, C.POINTER(_Matrix), C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.POINTER(_Matrix), C.POINTER(_vp)]),
        "p3g_tree_salt": (C.c_int, [_vp, C.c_uint64, _u64p, C.c_uint64, _u64p]),
        "p3g_tree_layer": (C.c_int, [_vp, C.c_uint32, _u64p]),
        "p3g_tree_matrix": (C.c_int, [_vp, C.c_uint64, C.POINTER(_vp), _u64p, _u64p]),
        "p3g_tree_prove_batch": (C.c_int, [_vp, _u64p, C.c_uint64, _vp]),
        "p3g_tree_free": (None, [_vp]),
        "p3g_commit_traces": (C.c_int, [_vp, C.c_uint32, C.POINTER(_Matrix), C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
        "p3g_server_set_batch_window": (C.c_int, [_vp, C.c_double]),
        "p3g_server_set_upload_ahead": (C.c_int, [_vp, C.c_uint32]),
        "p3g_quotient_extend_scale": (C.c_int, [_vp, _u64p, C.c_uint64, C.c_uint64, _u64p]),
        "p3g_quotient_divide_vanishing": (C.c_int, [_vp, _u64p, C.c_uint64, C.c_uint64, C.c_uint64]),
        "p3g_commit_quotient": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                          C.POINTER(_vp)]),
        "p3g_transcript_create": (C.c_int, [C.c_uint32, C.POINTER(_vp)]),
        "p3g_transcript_free": (None, [_vp]),
        "p3g_transcript_clone": (C.c_int, [_vp, C.POINTER(_vp)]),
        "p3g_transcript_observe": (C.c_int, [_vp, _u64p, C.c_uint64]),
        "p3g_transcript_observe_commitment": (C.c_int, [_vp, _u64p]),
        "p3g_transcript_sample": (C.c_int, [_vp, _u64p]),
        "p3g_transcript_sample_ext": (C.c_int, [_vp, _u64p]),
        "p3g_transcript_sample_bits": (C.c_int, [_vp, C.c_uint32, _u64p]),
        "p3g_transcript_send_fields": (C.c_int, [_vp, _u64p, C.c_uint64]),
        "p3g_transcript_send_commitment": (C.c_int, [_vp, _u64p]),
        "p3g_transcript_hint_fields": (C.c_int, [_vp, _u64p, C.c_uint64]),
        "p3g_transcript_hint_commitment": (C.c_int, [_vp, _u64p]),
        "p3g_transcript_grind": (C.c_int, [_vp, C.c_uint32, _u64p]),
        "p3g_transcript_num_fields": (C.c_uint64, [_vp]),
        "p3g_transcript_num_commitments": (C.c_uint64, [_vp]),
        "p3g_transcript_data": (C.c_int, [_vp, _u64p, _u64p]),
        "p3g_proof_encode": (C.c_int, [_u64p, C.c_uint64, _u64p, C.c_uint64, _vp, C.c_uint64, _u64p]),
        "p3g_proof_decode": (C.c_int, [_vp, C.c_uint64, _u64p, _u64p, _u64p, _u64p]),
        "p3g_transcript_encode": (C.c_int, [_vp, _vp, C.c_uint64, _u64p]),
        "p3g_open_with_channel": (C.c_int, [_vp, C.POINTER(PcsParams), C.c_uint32, _u64p, C.c_uint32, C.POINTER(_vp), C.c_uint64, _vp]),
        "p3g_deep_batch_eval": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint32, C.POINTER(_vp), C.c_uint64, _u64p,
                                          C.c_uint64, _u64p]),
        "p3g_deep_quotient": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint32, C.POINTER(_vp), C.c_uint64, _u64p,
                                        C.c_uint64, _u64p, _u64p, _u64p]),
        "p3g_fri_fold": (C.c_int, [_vp, _u64p, C.c_uint64, C.c_uint32, _u64p, _u64p]),
        "p3g_fri_final_poly": (C.c_int, [_u64p, C.c_uint64, _u64p]),
        "p3g_fri_commit": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint64, C.c_uint32, C.POINTER(_vp), _u64p]),
        "p3g_tree_prove_batch_cb": (C.c_int, [_vp, _u64p, C.c_uint64, C.POINTER(_Channel)]),
        "p3g_open_with_callbacks": (C.c_int, [_vp, C.POINTER(PcsParams), C.c_uint32, _u64p, C.c_uint32, C.POINTER(_vp), C.c_uint64, C.POINTER(_Channel)]),
        "p3g_commit_plan_create": (C.c_int, [_vp, C.c_uint32, _u64p, _u64p, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
        "p3g_commit_plan_input": (C.c_int, [_vp, C.c_uint64, C.POINTER(_vp)]),
        "p3g_commit_plan_info": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32)]),
        "p3g_commit_plan_run": (C.c_int, [_vp, C.POINTER(_vp)]),
        "p3g_commit_plan_destroy": (None, [_vp]),
        "p3g_prove_planned": (C.c_int, [_vp, _vp, C.POINTER(PcsParams), _vp, _u64p, _u64p]),
        "p3g_fill_random_device": (C.c_int, [_vp, C.c_uint64, C.c_uint64, _vp]),
        "p3g_request_footprint": (C.c_uint64, [C.POINTER(_Request)]),
        "p3g_server_create": (C.c_int, [C.c_int, C.c_uint32, C.c_uint64, C.c_double, C.c_uint32, C.POINTER(_vp)]),
        "p3g_server_submit": (C.c_int, [_vp, C.POINTER(_Request), _u64p]),
        "p3g_server_wait": (C.c_int, [_vp, C.c_uint64, C.POINTER(_ResultInfo)]),
        "p3g_server_poll": (C.c_int, [_vp, C.c_uint64, C.POINTER(C.c_int)]),
        "p3g_server_proof": (C.c_int, [_vp, C.c_uint64, _u64p, _u64p]),
        "p3g_server_cancel": (C.c_int, [_vp, C.c_uint64]),
        "p3g_server_release": (C.c_int, [_vp, C.c_uint64]),
        "p3g_server_stats": (C.c_int, [_vp, C.POINTER(_ServerCounters)]),
        "p3g_server_destroy": (None, [_vp]),
        "p3g_server_trim": (C.c_int, [_vp]),
        "p3g_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(_vp)]),
        "p3g_host_free": (C.c_int, [_vp]),
        "p3g_prove_many": (C.c_int, [_vp, C.c_uint32, C.POINTER(PcsParams), C.POINTER(_Matrix), C.c_uint64, C.c_uint64,
                                     C.c_uint32, C.POINTER(_vp), _u64p, _u64p]),
        "p3g_prove": (C.c_int, [_vp, C.c_uint32, C.POINTER(PcsParams), C.POINTER(_Matrix), C.c_uint64, C.c_uint32, _vp, _u64p, _u64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = _load()


def lib():
    if _lib is None:
        raise ImportError(f"libp3gpu.so not built at {LIB_PATH}; run __graft_entry__.build()")
    return _lib


def _check(rc: int, where: str) -> None:
    if rc != 0:
        raise P3GError(rc, where)


def u64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


def _p(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def fill_random(seed: int, n: int) -> np.ndarray:
    """RowMajorMatrix::rand values from SmallRng::seed_from_u64(seed), flat."""
    out = np.empty(n, np.uint64)
    _check(lib().p3g_fill_random(seed, n, _p(out)), "fill_random")
    return out


class Device:
    """One context per GPU: stream, constants, cached tables and device memory."""

    def __init__(self, device: int = 0, stream: int = None):
        """stream: optional hipStream_t handle that the context runs on instead of its own.  It must
        come from the HIP runtime this library links (libamdhip64.so.7 of the ROCm install; a runtime
        bundled elsewhere, such as inside a Python wheel, has its own handles) and outlive the Device."""
        h = _vp()
        if stream is None:
            _check(lib().p3g_ctx_create(device, C.byref(h)), "ctx_create")
        else:
            _check(lib().p3g_ctx_create_on_stream(device, _vp(stream), C.byref(h)), "ctx_create_on_stream")
        self.h = h
        self.index = device
        # trees, plans and matrices living in this context: freed before the context is
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)
        return obj

    def close(self):
        if getattr(self, "h", None):
            for c in list(self._children):
                try:
                    c.free()
                except Exception:
                    pass
            lib().p3g_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().p3g_ctx_synchronize(self.h), "synchronize")

    def trim(self):
        _check(lib().p3g_ctx_trim(self.h), "trim")

    def memory(self, reset_peak: bool = False) -> dict:
        """Device bytes: live allocations, block cache, peak live since the last reset."""
        v = np.zeros(3, np.uint64)
        _check(lib().p3g_ctx_memory(self.h, _p(v[0:1]), _p(v[1:2]), _p(v[2:3]), int(reset_peak)), "ctx_memory")
        return {"live": int(v[0]), "cached": int(v[1]), "peak_live": int(v[2])}

    def profile(self, on: bool = True):
        """Record HIP events around every launch of the named kernel families."""
        _check(lib().p3g_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_reset(self):
        _check(lib().p3g_profile_reset(self.h), "profile_reset")

    def profile_read(self, name: str):
        """(total milliseconds, launches) of one kernel family since the last reset."""
        ms = C.c_double(0)
        n = np.zeros(1, np.uint64)
        _check(lib().p3g_profile_read(self.h, name.encode(), C.byref(ms), _p(n)), "profile_read")
        return float(ms.value), int(n[0])

    VALU_KINDS = ("v_add_u32", "v_mad_u64_u32", "v_fma_f64", "v_mul_hi_u32", "add64_carry_pair", "v_mov_b32",
                  "v_cndmask_b32", "v_add_co_u32_e64", "v_lshl_add_u64", "v_add3_u32", "v_xor_b32", "v_alignbit_b32",
                  "v_cndmask_b32_e32", "v_mul_u32_u24_e32", "v_mad_u64_u32+v_add_u32_e32",
                  "v_mad_u64_u32+v_cndmask_b32_e32")

    def calibrate_valu(self, kind: int, iters: int = 16384, blocks: int = 8192) -> dict:
        """Sustained issue rate of one VALU instruction class (independent chains on every lane):
        launch time, instructions per wave, and wave64 instructions and lane-ops per second."""
        ms = C.c_double(0)
        n = np.zeros(1, np.uint64)
        _check(lib().p3g_calibrate_valu(self.h, kind, iters, blocks, C.byref(ms), _p(n)), "calibrate_valu")
        waves = blocks * 256 // 64
        insts = int(n[0]) * waves
        return {"kind": self.VALU_KINDS[kind], "ms": float(ms.value), "insts_per_wave": int(n[0]), "waves": waves,
                "wave_insts_per_s": insts / (ms.value * 1e-3), "lane_ops_per_s": 64 * insts / (ms.value * 1e-3)}

    def field_check(self, op: int, a, b, c=None, e: int = 0):
        """Run one of the kernels' device field operations on operand arrays a, b (and c):
        returns (out, wave_flags); see p3g_field_check in include/p3gpu.h."""
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.uint64)
        out = np.zeros(a.size, np.uint64)
        flags = np.zeros((a.size + 63) // 64, np.uint64)
        _check(lib().p3g_field_check(self.h, op, e, _p(a), _p(b), None if cc is None else _p(cc), a.size, _p(out),
                                     _p(flags)), "field_check")
        return out, flags

    def calibrate_copy(self, nbytes: int = 4 << 30, reps: int = 5) -> dict:
        """Device-to-device copy bandwidth (read + write bytes per second) of a copy kernel and of
        hipMemcpyAsync over nbytes."""
        k = C.c_double(0)
        m = C.c_double(0)
        _check(lib().p3g_calibrate_copy(self.h, nbytes, reps, C.byref(k), C.byref(m)), "calibrate_copy")
        return {"bytes": nbytes, "kernel_ms": float(k.value), "memcpy_ms": float(m.value),
                "kernel_gbs": 2 * nbytes / (k.value * 1e-3) / 1e9, "memcpy_gbs": 2 * nbytes / (m.value * 1e-3) / 1e9}

    def alloc_matrix(self, height: int, width: int) -> "DeviceMatrix":
        """Uninitialised height x width matrix in HBM."""
        ptr = _vp()
        _check(lib().p3g_alloc(self.h, max(8, 8 * height * width), C.byref(ptr)), "alloc")
        return self._adopt(DeviceMatrix(self, ptr, height, width))

    def upload(self, arr: np.ndarray) -> "DeviceMatrix":
        a = u64(arr)
        ptr = _vp()
        _check(lib().p3g_alloc(self.h, max(8, a.nbytes), C.byref(ptr)), "alloc")
        _check(lib().p3g_upload(self.h, ptr, a.ctypes.data_as(_vp), a.nbytes), "upload")
        shape = a.shape if a.ndim == 2 else (a.size, 1)
        return self._adopt(DeviceMatrix(self, ptr, shape[0], shape[1]))


class DeviceMatrix:
    """A row-major matrix resident in HBM (owned by this object unless handed to a tree)."""

    def __init__(self, dev: Device, ptr, height: int, width: int, owned: bool = True):
        self.dev, self.ptr, self.height, self.width, self.owned = dev, ptr, height, width, owned

    def view(self, offset_words: int, height: int, width: int) -> "DeviceMatrix":
        """A non-owning height x width matrix starting offset_words into this buffer."""
        if offset_words + height * width > self.height * self.width:
            raise ValueError("view out of range")
        base = self.ptr.value if isinstance(self.ptr, _vp) else int(self.ptr)
        return DeviceMatrix(self.dev, _vp(base + 8 * offset_words), height, width, owned=False)

    def fill_random(self, seed: int) -> None:
        """Canonical synthetic values from the device counter generator."""
        fill_random_device(self.dev, seed, self.height * self.width, self.ptr)

    def copy_from(self, other: "DeviceMatrix") -> None:
        """Device-to-device copy of a matrix of the same size."""
        if other.height * other.width != self.height * self.width:
            raise ValueError("size mismatch")
        _check(lib().p3g_copy_device(self.dev.h, self.ptr, other.ptr, 8 * self.height * self.width), "copy_device")

    def download(self) -> np.ndarray:
        out = np.empty(self.height * self.width, np.uint64)
        _check(lib().p3g_download(self.dev.h, out.ctypes.data_as(_vp), self.ptr, out.nbytes), "download")
        return out.reshape(self.height, self.width)

    def free(self):
        if self.owned and self.ptr and self.dev.h:
            _check(lib().p3g_free(self.dev.h, self.ptr), "free")
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _matrices(mats):
    """Build a p3g_matrix array; returns (array, keepalive, device_flag)."""
    arr = (_Matrix * max(1, len(mats)))()
    keep = []
    dev = None
    for i, m in enumerate(mats):
        if isinstance(m, DeviceMatrix):
            isdev = True
            arr[i] = _Matrix(C.cast(m.ptr, _u64p), m.height, m.width)
        else:
            isdev = False
            a = u64(m)
            if a.ndim != 2:
                raise ValueError("matrices must be 2-D")
            keep.append(a)
            arr[i] = _Matrix(_p(a), a.shape[0], a.shape[1])
        if dev is None:
            dev = isdev
        elif dev != isdev:
            raise ValueError("mix of host and device matrices")
    return arr, keep, bool(dev)


class Dft:
    """TwoAdicSubgroupDft on the GPU."""

    def __init__(self, device: Device):
        self.dev = device

    def coset_lde_batch(self, mat, added_bits: int, shift: int, bit_reverse_rows: bool = True) -> np.ndarray:
        a = u64(mat)
        n, w = a.shape
        out = np.empty((n << added_bits) * w, np.uint64)
        flags = BITREV_ROWS if bit_reverse_rows else 0
        _check(lib().p3g_coset_lde_batch(self.dev.h, _p(a), n, w, added_bits, shift, flags, _p(out)), "coset_lde_batch")
        return out.reshape(n << added_bits, w)


class ProverTranscript:
    """ProverTranscript over DuplexChallenger (Poseidon2) or the Keccak challenger."""

    def __init__(self, kind: int = CHALLENGER_DUPLEX_P2, _h=None):
        self.kind = kind
        if _h is None:
            _h = _vp()
            _check(lib().p3g_transcript_create(kind, C.byref(_h)), "transcript_create")
        self.h = _h

    def __del__(self):
        if getattr(self, "h", None):
            lib().p3g_transcript_free(self.h)
            self.h = None

    def clone(self) -> "ProverTranscript":
        h = _vp()
        _check(lib().p3g_transcript_clone(self.h, C.byref(h)), "transcript_clone")
        return ProverTranscript(self.kind, h)

    def observe(self, values):
        a = u64(np.atleast_1d(values))
        _check(lib().p3g_transcript_observe(self.h, _p(a), a.size), "observe")

    def observe_commitment(self, c):
        a = u64(c)
        _check(lib().p3g_transcript_observe_commitment(self.h, _p(a)), "observe_commitment")

    def sample(self) -> int:
        o = np.zeros(1, np.uint64)
        _check(lib().p3g_transcript_sample(self.h, _p(o)), "sample")
        return int(o[0])

    def sample_algebra_element(self):
        o = np.zeros(2, np.uint64)
        _check(lib().p3g_transcript_sample_ext(self.h, _p(o)), "sample_ext")
        return (int(o[0]), int(o[1]))

    def sample_bits(self, bits: int) -> int:
        o = np.zeros(1, np.uint64)
        _check(lib().p3g_transcript_sample_bits(self.h, bits, _p(o)), "sample_bits")
        return int(o[0])

    def send_field_slice(self, values):
        a = u64(values)
        _check(lib().p3g_transcript_send_fields(self.h, _p(a), a.size), "send_fields")

    def send_commitment(self, c):
        a = u64(c)
        _check(lib().p3g_transcript_send_commitment(self.h, _p(a)), "send_commitment")

    def hint_field_slice(self, values):
        a = u64(values)
        _check(lib().p3g_transcript_hint_fields(self.h, _p(a), a.size), "hint_fields")

    def hint_commitment(self, c):
        a = u64(c)
        _check(lib().p3g_transcript_hint_commitment(self.h, _p(a)), "hint_commitment")

    def grind(self, bits: int) -> int:
        o = np.zeros(1, np.uint64)
        _check(lib().p3g_transcript_grind(self.h, bits, _p(o)), "grind")
        return int(o[0])

    def encode(self) -> bytes:
        """The proof in the library's wire format (see encode_proof)."""
        n = np.zeros(1, np.uint64)
        _check(lib().p3g_transcript_encode(self.h, None, 0, _p(n)), "transcript_encode")
        buf = C.create_string_buffer(max(1, int(n[0])))
        _check(lib().p3g_transcript_encode(self.h, C.cast(buf, _vp), int(n[0]), _p(n)), "transcript_encode")
        return buf.raw[: int(n[0])]

    def data(self):
        """TranscriptData: (fields [n], commitments [m, 4])."""
        nf = lib().p3g_transcript_num_fields(self.h)
        nc = lib().p3g_transcript_num_commitments(self.h)
        # filled in place: a wide proof's fields run to megabytes, and every extra
        # pass over them shows in the proof rate
        f = np.empty(max(1, nf), np.uint64)
        c = np.empty(max(1, 4 * nc), np.uint64)
        _check(lib().p3g_transcript_data(self.h, _p(f), _p(c)), "transcript_data")
        return f[:nf], c[: 4 * nc].reshape(nc, 4)


def proof_bytes(fields: np.ndarray, commitments: np.ndarray) -> bytes:
    """The proof's words without a header (field stream, then commitment stream, each u64 LE): the
    input of proof digests.  encode_proof is the self-describing wire format."""
    return u64(fields).astype("<u8").tobytes() + u64(commitments).ravel().astype("<u8").tobytes()


class LiftedMerkleTree:
    def __init__(self, h, hash_kind: int, dev: "Device"):
        self.h = h
        self.hash_kind = hash_kind
        self.dev = dev
        dev._adopt(self)

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "h", None):
            if self.dev.h:
                lib().p3g_tree_free(self.h)
            self.h = None

    def root(self) -> np.ndarray:
        o = np.zeros(4, np.uint64)
        _check(lib().p3g_tree_root(self.h, _p(o)), "tree_root")
        return o

    def height(self) -> int:
        return int(lib().p3g_tree_height(self.h))

    def alignment(self) -> int:
        return int(lib().p3g_tree_alignment(self.h))

    def widths(self):
        n = lib().p3g_tree_num_matrices(self.h)
        o = np.zeros(max(1, n), np.uint64)
        _check(lib().p3g_tree_widths(self.h, _p(o)), "tree_widths")
        return [int(x) for x in o[:n]]

    def rows(self, index: int) -> np.ndarray:
        ln = np.zeros(1, np.uint64)
        _check(lib().p3g_tree_rows(self.h, index, None, 0, _p(ln)), "tree_rows")
        o = np.zeros(max(1, int(ln[0])), np.uint64)
        _check(lib().p3g_tree_rows(self.h, index, _p(o), o.size, _p(ln)), "tree_rows")
        return o[: int(ln[0])]

    def salt(self, index: int) -> np.ndarray:
        """LiftedMerkleTree::salt(index): the leaf's salt (empty for a tree without salt)."""
        n = np.zeros(1, np.uint64)
        _check(lib().p3g_tree_salt(self.h, index, None, 0, _p(n)), "tree_salt")
        o = np.zeros(max(1, int(n[0])), np.uint64)
        if int(n[0]):
            _check(lib().p3g_tree_salt(self.h, index, _p(o), o.size, _p(n)), "tree_salt")
        return o[: int(n[0])]

    def layer(self, l: int) -> np.ndarray:
        n = self.height() >> l
        o = np.zeros(4 * n, np.uint64)
        _check(lib().p3g_tree_layer(self.h, l, _p(o)), "tree_layer")
        return o.reshape(n, 4)

    def matrix(self, i: int):
        ptr = _vp()
        h = np.zeros(1, np.uint64)
        w = np.zeros(1, np.uint64)
        _check(lib().p3g_tree_matrix(self.h, i, C.byref(ptr), _p(h), _p(w)), "tree_matrix")
        return ptr, int(h[0]), int(w[0])

    def prove_batch(self, indices, channel):
        a = u64(list(indices))
        if isinstance(channel, ProverTranscript):
            _check(lib().p3g_tree_prove_batch(self.h, _p(a) if a.size else None, a.size, channel.h), "prove_batch")
        else:
            cb = channel if isinstance(channel, CallbackChannel) else CallbackChannel(channel)
            _check(lib().p3g_tree_prove_batch_cb(self.h, _p(a) if a.size else None, a.size, C.byref(cb.struct)),
                   "prove_batch_cb")


class LmcsConfig:
    """LmcsConfig over Poseidon2-Goldilocks-12 or the u64-serialising Keccak sponge."""

    def __init__(self, device: Device, hash_kind: int = HASH_POSEIDON2):
        self.dev = device
        self.hash_kind = hash_kind

    def alignment(self) -> int:
        return int(lib().p3g_lmcs_alignment(self.hash_kind))

    def _build(self, mats, alignment, take_ownership=False, salt=None):
        arr, keep, dev = _matrices(mats)
        flags = (DEVICE_PTRS if dev else 0) | (TAKE_OWNERSHIP if (dev and take_ownership) else 0)
        h = _vp()
        if salt is None:
            _check(lib().p3g_lmcs_build(self.dev.h, self.hash_kind, arr, len(mats), alignment, flags, C.byref(h)),
                   "lmcs_build")
        else:
            sarr, skeep, sdev = _matrices([salt])
            if sdev != dev:
                raise ValueError("salt and matrices must both be host arrays or both device matrices")
            _check(lib().p3g_lmcs_build_salted(self.dev.h, self.hash_kind, arr, len(mats), alignment, flags, sarr,
                                               C.byref(h)), "lmcs_build_salted")
        if dev and take_ownership:
            for m in list(mats) + ([salt] if salt is not None else []):
                m.owned = False
        return LiftedMerkleTree(h, self.hash_kind, self.dev)

    def build_tree(self, mats, take_ownership=False, salt=None) -> LiftedMerkleTree:
        """build_tree; with salt (a tree-height x SALT matrix), the hiding form of
        HidingLmcsConfig::build_tree (hiding_lmcs.rs:104-118) with the caller's salt."""
        return self._build(mats, 1, take_ownership, salt)

    def build_aligned_tree(self, mats, take_ownership=False, salt=None) -> LiftedMerkleTree:
        return self._build(mats, self.alignment(), take_ownership, salt)

    def hash(self, rows) -> np.ndarray:
        rows = [u64(r).ravel() for r in rows]
        flat = u64(np.concatenate(rows)) if rows else np.zeros(1, np.uint64)
        ws = u64([r.size for r in rows]) if rows else np.zeros(1, np.uint64)
        o = np.zeros(4, np.uint64)
        _check(lib().p3g_lmcs_hash(self.hash_kind, _p(flat), _p(ws), len(rows), _p(o)), "lmcs_hash")
        return o

    def compress(self, left, right) -> np.ndarray:
        a, b, o = u64(left), u64(right), np.zeros(4, np.uint64)
        _check(lib().p3g_lmcs_compress(self.hash_kind, _p(a), _p(b), _p(o)), "lmcs_compress")
        return o


def commit_traces(lmcs: LmcsConfig, traces, log_blowup: int) -> LiftedMerkleTree:
    arr, keep, dev = _matrices(traces)
    h = _vp()
    _check(lib().p3g_commit_traces(lmcs.dev.h, lmcs.hash_kind, arr, len(traces), log_blowup,
                                   DEVICE_PTRS if dev else 0, C.byref(h)), "commit_traces")
    return LiftedMerkleTree(h, lmcs.hash_kind, lmcs.dev)


def encode_proof(fields, commitments) -> bytes:
    """TranscriptData (data.rs:8-14) as bytes: "P3GT", version 1 (u32 LE), the field count and
    the commitment count (u64 LE each), then every canonical word as u64 LE."""
    f = u64(np.asarray(fields, dtype=np.uint64).ravel())
    c = u64(np.asarray(commitments, dtype=np.uint64).ravel())
    if c.size % 4:
        raise ValueError("commitments are 4 words each")
    n = np.zeros(1, np.uint64)
    fp, cp = (_p(f) if f.size else None), (_p(c) if c.size else None)
    _check(lib().p3g_proof_encode(fp, f.size, cp, c.size // 4, None, 0, _p(n)), "proof_encode")
    buf = C.create_string_buffer(max(1, int(n[0])))
    _check(lib().p3g_proof_encode(fp, f.size, cp, c.size // 4, C.cast(buf, _vp), int(n[0]), _p(n)), "proof_encode")
    return buf.raw[: int(n[0])]


def decode_proof(data: bytes):
    """Inverse of encode_proof: (fields [n], commitments [m, 4])."""
    buf = C.create_string_buffer(bytes(data), max(1, len(data)))
    nf, nc = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
    _check(lib().p3g_proof_decode(C.cast(buf, _vp), len(data), None, _p(nf), None, _p(nc)), "proof_decode")
    f = np.zeros(max(1, int(nf[0])), np.uint64)
    c = np.zeros(max(1, 4 * int(nc[0])), np.uint64)
    _check(lib().p3g_proof_decode(C.cast(buf, _vp), len(data), _p(f), _p(nf), _p(c), _p(nc)), "proof_decode")
    return f[: int(nf[0])].copy(), c[: 4 * int(nc[0])].reshape(int(nc[0]), 4).copy()


def commit_aux_traces(lmcs: LmcsConfig, aux_traces, log_blowup: int) -> LiftedMerkleTree:
    """The auxiliary-trace commitment of prove_multi (p3-miden-lifted-stark/src/prover/mod.rs:323-337):
    each extension-field trace is flattened to base (c0 then c1 per element, width 2w) and the
    flattened traces go through commit_traces.  Host traces are (n, w, 2) arrays; a DeviceMatrix
    already holds the flattened n x 2w words."""
    flat = []
    for t in aux_traces:
        if isinstance(t, DeviceMatrix):
            flat.append(t)
        else:
            a = u64(t)
            if a.ndim != 3 or a.shape[2] != 2:
                raise ValueError("host aux traces are (n, w, 2) arrays of extension elements")
            flat.append(a.reshape(a.shape[0], 2 * a.shape[1]))
    return commit_traces(lmcs, flat, log_blowup)


def cyclic_extend_and_scale(acc: "DeviceMatrix", m: int, target: int, beta) -> None:
    """cyclic_extend_and_scale (p3-miden-lifted-stark/src/prover/quotient.rs:40-55) in place on a
    device buffer of at least `target` extension elements (rows of 2 words) whose first m hold the
    accumulator; m == 0 zero-fills."""
    if acc.height * acc.width < 2 * target:
        raise ValueError("buffer smaller than target extension elements")
    b = u64(np.asarray(beta, dtype=np.uint64).ravel())
    _check(lib().p3g_quotient_extend_scale(acc.dev.h, C.cast(acc.ptr, _u64p), m, target, _p(b)),
           "quotient_extend_scale")


def divide_by_vanishing(num: "DeviceMatrix", n: int, d: int, shift: int) -> None:
    """divide_by_vanishing_in_place (quotient.rs:72-101) in place on n*d extension elements of the
    quotient coset shift*J, natural order (device buffer, rows of 2 words)."""
    if num.height * num.width < 2 * n * d:
        raise ValueError("buffer smaller than n*d extension elements")
    _check(lib().p3g_quotient_divide_vanishing(num.dev.h, C.cast(num.ptr, _u64p), n, d, shift),
           "quotient_divide_vanishing")


def commit_quotient(lmcs: LmcsConfig, q_evals, d: int, log_blowup: int) -> LiftedMerkleTree:
    """commit_quotient (p3-miden-lifted-stark/src/prover/quotient.rs:133-238).

    q_evals: the quotient's N*D extension elements in natural order on gJ, as an
    (N*D, 2) host array of (c0, c1) or a DeviceMatrix of N rows x 2D words.  The tree
    holds the bit-reversed (N*B) x 2D LDE of the D chunk polynomials on gK."""
    if isinstance(q_evals, DeviceMatrix):
        if q_evals.width != 2 * d:
            raise ValueError("device quotient matrix must be N x 2D")
        ptr, n, flags, keep = C.cast(q_evals.ptr, _u64p), q_evals.height, DEVICE_PTRS, None
    else:
        keep = u64(q_evals).reshape(-1)
        if d <= 0 or keep.size % (2 * d):
            raise ValueError("q_evals length must be a multiple of 2D words")
        ptr, n, flags = _p(keep), keep.size // (2 * d), 0
    h = _vp()
    _check(lib().p3g_commit_quotient(lmcs.dev.h, lmcs.hash_kind, ptr, n, d, log_blowup, flags, C.byref(h)),
           "commit_quotient")
    return LiftedMerkleTree(h, lmcs.hash_kind, lmcs.dev)


def open_with_channel(params: PcsParams, lmcs: LmcsConfig, log_lde_height: int, eval_points, trace_trees,
                      channel) -> None:
    """Runs on the library's own transcript when given a ProverTranscript; any other
    channel object is driven through callbacks (CallbackChannel)."""
    z = u64(np.asarray(eval_points, dtype=np.uint64).ravel())
    hs = (_vp * len(trace_trees))(*[t.h for t in trace_trees])
    if isinstance(channel, ProverTranscript):
        _check(lib().p3g_open_with_channel(lmcs.dev.h, C.byref(params), log_lde_height, _p(z), z.size // 2, hs,
                                           len(trace_trees), channel.h), "open_with_channel")
    else:
        cb = channel if isinstance(channel, CallbackChannel) else CallbackChannel(channel)
        _check(lib().p3g_open_with_callbacks(lmcs.dev.h, C.byref(params), log_lde_height, _p(z), z.size // 2, hs,
                                             len(trace_trees), C.byref(cb.struct)), "open_with_callbacks")


def batch_eval_lifted(trace_trees, eval_points, log_blowup: int) -> np.ndarray:
    """PointQuotients::batch_eval_lifted (p3-miden-lifted-fri/src/deep/interpolate.rs:123-201) of every
    committed column of the trees at the points: an array [column][point][2] (unaligned widths, tree
    then matrix order)."""
    if not trace_trees:
        raise ValueError("at least one tree required")
    z = u64(np.asarray(eval_points, dtype=np.uint64).ravel())
    hs = (_vp * len(trace_trees))(*[t.h for t in trace_trees])
    h = trace_trees[0].dev.h
    n = np.zeros(1, np.uint64)
    _check(lib().p3g_deep_batch_eval(h, log_blowup, _p(z), z.size // 2, hs, len(trace_trees), None, 0, _p(n)),
           "deep_batch_eval")
    out = np.zeros(max(1, int(n[0])), np.uint64)
    _check(lib().p3g_deep_batch_eval(h, log_blowup, _p(z), z.size // 2, hs, len(trace_trees), _p(out), out.size,
                                     _p(n)), "deep_batch_eval")
    return out[: int(n[0])].reshape(-1, z.size // 2, 2)


def deep_quotient(trace_trees, eval_points, log_blowup: int, evals, alpha, beta) -> "DeviceMatrix":
    """DeepPoly::from_evals after its challenges (p3-miden-lifted-fri/src/deep/prover.rs:106-285): the DEEP
    polynomial's L evaluations (bit-reversed LDE order) as an L x 2 DeviceMatrix, from `evals` as
    batch_eval_lifted returns them and the challenges alpha (columns) and beta (points)."""
    if not trace_trees:
        raise ValueError("at least one tree required")
    z = u64(np.asarray(eval_points, dtype=np.uint64).ravel())
    ev = u64(np.asarray(evals, dtype=np.uint64).ravel())
    a = u64(np.asarray(alpha, dtype=np.uint64).ravel())
    b = u64(np.asarray(beta, dtype=np.uint64).ravel())
    hs = (_vp * len(trace_trees))(*[t.h for t in trace_trees])
    dev = trace_trees[0].dev
    out = dev.alloc_matrix(trace_trees[0].height(), 2)
    _check(lib().p3g_deep_quotient(dev.h, log_blowup, _p(z), z.size // 2, hs, len(trace_trees), _p(ev), ev.size,
                                   _p(a), _p(b), C.cast(out.ptr, _u64p)), "deep_quotient")
    return out


def fri_fold(evals: "DeviceMatrix", log_arity: int, beta) -> "DeviceMatrix":
    """One FRI round's fold (FriFold::fold_matrix, p3-miden-lifted-fri/src/fri/fold/mod.rs:119-144) of the
    n bit-reversed evaluations in an n x 2 DeviceMatrix into an (n >> log_arity) x 2 one."""
    if evals.width != 2:
        raise ValueError("evaluations must be an n x 2 matrix of extension elements")
    b = u64(np.asarray(beta, dtype=np.uint64).ravel())
    out = evals.dev.alloc_matrix(max(1, evals.height >> log_arity), 2)
    _check(lib().p3g_fri_fold(evals.dev.h, C.cast(evals.ptr, _u64p), evals.height, log_arity, _p(b),
                              C.cast(out.ptr, _u64p)), "fri_fold")
    return out


def fri_commit(lmcs: "LmcsConfig", evals: "DeviceMatrix", log_arity: int) -> "LiftedMerkleTree":
    """A FRI round's commitment (p3-miden-lifted-fri/src/fri/prover.rs:136-150): the unaligned tree over
    the n evaluations of an n x 2 DeviceMatrix read as the (n >> log_arity) x (2 << log_arity) base
    matrix.  The tree keeps a reference to `evals`."""
    if evals.width != 2:
        raise ValueError("evaluations must be an n x 2 matrix of extension elements")
    h = _vp()
    _check(lib().p3g_fri_commit(evals.dev.h, lmcs.hash_kind, C.cast(evals.ptr, _u64p), evals.height, log_arity,
                                C.byref(h), None), "fri_commit")
    t = LiftedMerkleTree(h, lmcs.hash_kind, evals.dev)
    t._evals = evals
    return t


def fri_final_poly(evals) -> np.ndarray:
    """The final polynomial (p3-miden-lifted-fri/src/fri/prover.rs:208-221): n evaluations of the last
    domain in bit-reversed order ([n][2] host array) to n coefficients in descending degree."""
    v = u64(np.asarray(evals, dtype=np.uint64).ravel())
    out = np.zeros(max(2, v.size), np.uint64)
    _check(lib().p3g_fri_final_poly(_p(v), v.size // 2, _p(out)), "fri_final_poly")
    return out[: v.size].reshape(-1, 2)


def prove(lmcs: LmcsConfig, params: PcsParams, traces, channel: ProverTranscript = None):
    """pcs_trace flow: commit_traces, observe root, sample two points, open.
    Returns (root, eval_points, fields, commitments)."""
    if channel is None:
        channel = ProverTranscript(CHALLENGER_DUPLEX_P2 if lmcs.hash_kind == HASH_POSEIDON2 else CHALLENGER_KECCAK)
    arr, keep, dev = _matrices(traces)
    root = np.zeros(4, np.uint64)
    z = np.zeros(4, np.uint64)
    _check(lib().p3g_prove(lmcs.dev.h, lmcs.hash_kind, C.byref(params), arr, len(traces),
                           DEVICE_PTRS if dev else 0, channel.h, _p(root), _p(z)), "prove")
    fields, comms = channel.data()
    return root, ((int(z[0]), int(z[1])), (int(z[2]), int(z[3]))), fields, comms


PLAN_GRAPH = 1
PLAN_AUTOTUNE = 2
PLAN_PER_TRACE = 4


class CommitPlan:
    """Batched commit of many traces of fixed shapes: grouped NTT launches over
    all traces, optional hipGraph replay and autotuned launch shapes.  Results
    are identical to commit_traces on the same traces."""

    def __init__(self, lmcs: LmcsConfig, shapes, log_blowup: int, graph: bool = True, autotune: bool = False,
                 per_trace: bool = False):
        self.lmcs = lmcs
        self.shapes = [(int(h), int(w)) for h, w in shapes]
        self.log_blowup = log_blowup
        hs = u64([h for h, _ in self.shapes]) if shapes else np.zeros(1, np.uint64)
        ws = u64([w for _, w in self.shapes]) if shapes else np.zeros(1, np.uint64)
        h = _vp()
        flags = (PLAN_GRAPH if graph else 0) | (PLAN_AUTOTUNE if autotune else 0) | (PLAN_PER_TRACE if per_trace else 0)
        _check(lib().p3g_commit_plan_create(lmcs.dev.h, lmcs.hash_kind, _p(hs), _p(ws), len(self.shapes), log_blowup,
                                            flags, C.byref(h)), "commit_plan_create")
        self.h = h
        lmcs.dev._adopt(self)

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "h", None):
            if self.lmcs.dev.h:
                lib().p3g_commit_plan_destroy(self.h)
            self.h = None

    def input_ptr(self, i: int) -> int:
        ptr = _vp()
        _check(lib().p3g_commit_plan_input(self.h, i, C.byref(ptr)), "commit_plan_input")
        return ptr.value

    def set_input(self, i: int, mat: np.ndarray) -> None:
        a = u64(mat)
        assert a.shape == self.shapes[i], (a.shape, self.shapes[i])
        _check(lib().p3g_upload(self.lmcs.dev.h, self.input_ptr(i), _p(a), a.size * 8), "upload")

    def fill_random(self, seed: int) -> None:
        for i, (h, w) in enumerate(self.shapes):
            _check(lib().p3g_fill_random_device(self.lmcs.dev.h, seed + i, h * w, self.input_ptr(i)),
                   "fill_random_device")

    def info(self):
        k, t, n, g = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(lib().p3g_commit_plan_info(self.h, C.byref(k), C.byref(t), C.byref(n), C.byref(g)), "commit_plan_info")
        return {"kmax": k.value, "leaf_threads": t.value, "launches": n.value, "grouped": bool(g.value)}

    def run(self) -> LiftedMerkleTree:
        """Recompute LDEs and tree; the returned view is valid until the next run."""
        t = _vp()
        _check(lib().p3g_commit_plan_run(self.h, C.byref(t)), "commit_plan_run")
        return LiftedMerkleTree(t, self.lmcs.hash_kind, self.lmcs.dev)

    def prove(self, params: PcsParams, channel: ProverTranscript = None):
        if channel is None:
            channel = ProverTranscript(CHALLENGER_DUPLEX_P2 if self.lmcs.hash_kind == HASH_POSEIDON2 else CHALLENGER_KECCAK)
        root = np.zeros(4, np.uint64)
        z = np.zeros(4, np.uint64)
        _check(lib().p3g_prove_planned(self.lmcs.dev.h, self.h, C.byref(params), channel.h, _p(root), _p(z)),
               "prove_planned")
        fields, comms = channel.data()
        return root, ((int(z[0]), int(z[1])), (int(z[2]), int(z[3]))), fields, comms


def prove_many(lmcs: LmcsConfig, params: PcsParams, jobs, channels=None):
    """Independent proofs of many same-shaped jobs (each a list of traces) in lock-step.
    Returns one (root, eval_points, fields, commitments) per job, each identical to prove()."""
    k = len(jobs)
    if k == 0:
        return []
    t = len(jobs[0])
    if any(len(j) != t for j in jobs):
        raise ValueError("every job needs the same number of traces")
    if channels is None:
        kind = CHALLENGER_DUPLEX_P2 if lmcs.hash_kind == HASH_POSEIDON2 else CHALLENGER_KECCAK
        channels = [ProverTranscript(kind) for _ in range(k)]
    flat = [m for j in jobs for m in j]
    arr, keep, dev = _matrices(flat)
    hs = (_vp * k)(*[c.h for c in channels])
    roots = np.zeros(4 * k, np.uint64)
    zs = np.zeros(4 * k, np.uint64)
    _check(lib().p3g_prove_many(lmcs.dev.h, lmcs.hash_kind, C.byref(params), arr, k, t, DEVICE_PTRS if dev else 0,
                                hs, _p(roots), _p(zs)), "prove_many")
    out = []
    for j, c in enumerate(channels):
        z = zs[4 * j: 4 * j + 4]
        f, cm = c.data()
        out.append((roots[4 * j: 4 * j + 4].copy(), ((int(z[0]), int(z[1])), (int(z[2]), int(z[3]))), f, cm))
    return out


def fill_random_device(device: Device, seed: int, n: int, ptr) -> None:
    _check(lib().p3g_fill_random_device(device.h, seed, n, ptr), "fill_random_device")


INJECT_ALLOC_FAILURE = 8  # request flag: failure-handling test hook (the request's third device allocation fails)


def _request(hash_kind, params: PcsParams, traces, extra_flags: int = 0):
    arr, keep, dev = _matrices(traces)
    req = _Request(hash_kind, params, C.cast(arr, C.POINTER(_Matrix)), len(traces),
                   (DEVICE_PTRS if dev else 0) | extra_flags)
    return req, (arr, keep, traces)


def request_footprint(hash_kind: int, params: PcsParams, traces) -> int:
    """Device bytes the server reserves for this request at admission."""
    req, _keep = _request(hash_kind, params, traces)
    return int(lib().p3g_request_footprint(C.byref(req)))


class ServedProof:
    def __init__(self, info, fields, comms):
        self.root = np.array(list(info.root), np.uint64)
        z = list(info.eval_points)
        self.eval_points = ((int(z[0]), int(z[1])), (int(z[2]), int(z[3])))
        self.fields, self.commitments = fields, comms
        self.submit_ms, self.admit_ms, self.done_ms = info.submit_ms, info.admit_ms, info.done_ms
        self.latency_ms = info.done_ms - info.submit_ms
        self.worker = int(info.worker)
        self.footprint_bytes = int(info.footprint_bytes)


class Server:
    """Prove-request server: W worker contexts on one GPU, admission by device
    footprint within an HBM budget, continuous batching of arriving requests."""

    def __init__(self, device: int = 0, workers: int = 4, hbm_budget_bytes: int = 250 * 10**9,
                 backfill_ms: float = 50.0, max_batch: int = 64, batch_window_ms: float = 0.0,
                 upload_ahead: int = 0):
        """upload_ahead: how many queued requests' host traces a stage of their own copies to the
        device ahead of admission, overlapping PCIe copies with proving (0: off)."""
        h = _vp()
        _check(lib().p3g_server_create(device, workers, int(hbm_budget_bytes), float(backfill_ms), int(max_batch),
                                       C.byref(h)), "server_create")
        self.h = h
        self._keep = {}
        if batch_window_ms:
            _check(lib().p3g_server_set_batch_window(h, float(batch_window_ms)), "server_set_batch_window")
        if upload_ahead:
            _check(lib().p3g_server_set_upload_ahead(h, int(upload_ahead)), "server_set_upload_ahead")

    def submit(self, traces, params: PcsParams, hash_kind: int = HASH_POSEIDON2, inject_alloc_failure=False) -> int:
        """Queue a proof of `traces` (host arrays or DeviceMatrix); returns a ticket.
        The arrays are kept alive here until the result is collected.  inject_alloc_failure:
        test hook making one of the request's device allocations fail (P3G_E_NOMEM)."""
        req, keep = _request(hash_kind, params, traces, INJECT_ALLOC_FAILURE if inject_alloc_failure else 0)
        t = np.zeros(1, np.uint64)
        _check(lib().p3g_server_submit(self.h, C.byref(req), _p(t)), "server_submit")
        self._keep[int(t[0])] = keep
        return int(t[0])

    def done(self, ticket: int) -> bool:
        d = C.c_int(0)
        _check(lib().p3g_server_poll(self.h, ticket, C.byref(d)), "server_poll")
        return bool(d.value)

    def result(self, ticket: int) -> ServedProof:
        """Wait for the request, return its proof and timings, and release it."""
        info = _ResultInfo()
        rc = lib().p3g_server_wait(self.h, ticket, C.byref(info))
        try:
            _check(rc, "server_wait")
            f = np.zeros(max(1, info.num_fields), np.uint64)
            c = np.zeros(max(1, 4 * info.num_commitments), np.uint64)
            _check(lib().p3g_server_proof(self.h, ticket, _p(f), _p(c)), "server_proof")
            return ServedProof(info, f[: info.num_fields], c[: 4 * info.num_commitments].reshape(-1, 4))
        finally:
            lib().p3g_server_release(self.h, ticket)
            self._keep.pop(ticket, None)

    def cancel(self, ticket: int):
        """Cancel a request: queued, it completes at once (or, while the upload stage is enqueuing
        its copies, as soon as they are enqueued); running, at its next phase boundary.
        Its result() then raises P3GError with code 9 (P3G_E_CANCELED), unless it had already finished."""
        _check(lib().p3g_server_cancel(self.h, ticket), "server_cancel")

    def stats(self) -> dict:
        s = _ServerCounters()
        _check(lib().p3g_server_stats(self.h, C.byref(s)), "server_stats")
        return {n: getattr(s, n) for n, _ in _ServerCounters._fields_}

    def trim(self):
        """Free the device blocks the server's pool holds but does not use."""
        _check(lib().p3g_server_trim(self.h), "server_trim")

    def close(self):
        if getattr(self, "h", None):
            lib().p3g_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        self.close()


class PinnedArray:
    """A numpy view of pinned host memory (p3g_host_alloc), for fast host->device copies."""

    def __init__(self, shape):
        n = int(np.prod(shape))
        p = _vp()
        _check(lib().p3g_host_alloc(max(8, 8 * n), C.byref(p)), "host_alloc")
        self.ptr = p
        buf = (C.c_uint64 * max(1, n)).from_address(p.value)
        self.array = np.ctypeslib.as_array(buf)[:n].reshape(shape)

    def free(self):
        if getattr(self, "ptr", None):
            self.array = None
            lib().p3g_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


from .air import AIR_OP_DTYPE, AirProgram  # noqa: E402


def _air_ops(program) -> np.ndarray:
    ops = program.ops if isinstance(program, AirProgram) else np.asarray(program)
    if ops.dtype != AIR_OP_DTYPE:
        raise TypeError("a constraint program is an AirProgram or an array of AIR_OP_DTYPE")
    return np.ascontiguousarray(ops)


def air_compile(program) -> int:
    """Check a constraint program and compile its kernel into the process cache ahead of
    eval_constraints (needs no GPU); returns the size of the gfx950 code object."""
    ops = _air_ops(program)
    size = C.c_uint64(0)
    _check(lib().p3g_air_compile(ops.ctypes.data_as(_vp), ops.size, C.byref(size)), "air_compile")
    return size.value


def eval_constraints(program, main: "DeviceMatrix", n: int, log_d: int, shift: int, alpha, accumulator: "DeviceMatrix",
                     aux: "DeviceMatrix" = None, periodic=None, publics=(), ext_inputs=()) -> None:
    """evaluate_constraints_into (p3-miden-lifted-stark/src/prover/constraints/mod.rs:63-179) on the GPU:
    accumulator (n*2^log_d extension elements, an (n*d) x 2 DeviceMatrix) += the program's constraints
    folded with powers of alpha, at every point of the quotient domain shift*<w_{n d}>.  main / aux are
    committed LDEs with bit-reversed rows (a tree's matrix, or any device matrix of that layout); periodic
    is the host table of the periodic columns' LDE on gJ (natural order, PeriodicLde)."""
    ops = _air_ops(program)
    if accumulator.height * accumulator.width < 2 * (n << log_d):
        raise ValueError("accumulator smaller than n*d extension elements")
    dev = main.dev
    mm = _Matrix(C.cast(main.ptr, _u64p), main.height, main.width)
    am = _Matrix(C.cast(aux.ptr, _u64p), aux.height, aux.width) if aux is not None else None
    per = u64(periodic) if periodic is not None else None
    pub = u64(np.atleast_1d(np.asarray(publics, dtype=np.uint64)))
    ext = u64(np.asarray(ext_inputs, dtype=np.uint64).reshape(-1))
    al = u64(alpha)
    _check(lib().p3g_eval_constraints(dev.h, ops.ctypes.data_as(_vp), ops.size, C.byref(mm),
                                      C.byref(am) if am is not None else None,
                                      _p(per) if per is not None else None, per.shape[0] if per is not None else 0,
                                      per.shape[1] if per is not None else 0, _p(pub), len(publics), _p(ext),
                                      ext.size // 2, n, log_d, shift, _p(al), accumulator.ptr), "eval_constraints")
