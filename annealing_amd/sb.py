"""ctypes binding of the solver's C ABI (include/sb.h).

This is the host-side mirror of the reference's solver boundary
(SB/CUDA_GPU/m2_opt.cu:100-220): load an instance, solve R replicas, read
energies and spins.  All compute runs in libsbamd.so (HIP kernels for
gfx950); this module only marshals numpy buffers.  There is no fallback: if
the library is missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsbamd.so")

ASB, BSB, DSB, ASB_M1 = 0, 1, 2, 3
PREC_FP32, PREC_SPLIT, PREC_ORDERED = 0, 1, 2
INIT_UNIFORM, INIT_REFERENCE = 0, 1
J_FP32, J_BF16, J_FP8 = 0, 1, 2
GEN_ER, GEN_SK, GEN_KN, GEN_SK_DILUTED = 0, 1, 2, 3

VARIANTS = {"asb": ASB, "bsb": BSB, "dsb": DSB, "asb_m1": ASB_M1}
PRECISIONS = {"fp32": PREC_FP32, "split": PREC_SPLIT, "ordered": PREC_ORDERED}
INITS = {"uniform": INIT_UNIFORM, "reference": INIT_REFERENCE}
SERVE_SHORTEST, SERVE_ARRIVAL = 0, 1
SERVE_POLICIES = {"shortest": SERVE_SHORTEST, "arrival": SERVE_ARRIVAL}

# Every symbol include/sb.h declares.
EXPORTS = (
    "sb_version", "sb_status_string", "sb_last_error_message", "sb_default_params",
    "sb_edgelist_parse", "sb_edgelist_parse_field", "sb_instance_stats", "sb_c0_auto", "sb_generate",
    "sb_ctx_create", "sb_ctx_destroy", "sb_problem_create",
    "sb_problem_from_edgelist_file", "sb_problem_generate", "sb_problem_read_j",
    "sb_problem_destroy", "sb_problem_info", "sb_problem_set_field", "sb_problem_set_sparse",
    "sb_problem_sparse_nnz",
    "sb_solve", "sb_step_coeffs_for", "sb_step", "sb_field", "sb_init_state", "sb_asb_prologue",
    "sb_energy", "sb_energy_noalloc", "sb_batch_solve", "sb_server_create", "sb_server_submit",
    "sb_server_submit_generated", "sb_server_wait", "sb_server_cancel", "sb_server_stats_get",
    "sb_server_destroy",
)


class SBError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        msg = f"{where}: status {status}"
        if _lib is not None:
            msg += f" ({_lib.sb_status_string(status).decode()})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


class Params(C.Structure):
    _fields_ = [
        ("variant", C.c_int32), ("precision", C.c_int32), ("init", C.c_int32),
        ("schedule", C.c_int32), ("steps", C.c_int32), ("replicas", C.c_int32),
        ("M", C.c_int32), ("reserved", C.c_int32), ("dt", C.c_double),
        ("a0", C.c_double), ("c0", C.c_double), ("K", C.c_double), ("seed", C.c_uint64),
    ]


class Result(C.Structure):
    _fields_ = [
        ("energy_i64", C.POINTER(C.c_int64)), ("energy_f64", C.POINTER(C.c_double)),
        ("spins", C.POINTER(C.c_int8)), ("x_final", C.POINTER(C.c_float)),
        ("y_final", C.POINTER(C.c_float)), ("best_replica", C.c_int32),
        ("energy_exact", C.c_int32), ("c0_used", C.c_double), ("step_ms", C.c_double),
        ("total_ms", C.c_double), ("cut", C.POINTER(C.c_double)),
    ]


class StepCoeffs(C.Structure):
    _fields_ = [
        ("variant", C.c_int32), ("do_local", C.c_int32), ("M", C.c_int32),
        ("reserved", C.c_int32), ("alpha", C.c_float), ("beta", C.c_float),
        ("gamma", C.c_float), ("p_next", C.c_float), ("a0", C.c_double),
        ("K", C.c_double), ("dts", C.c_double), ("c0", C.c_double),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("work_items", C.c_int32), ("workgroups", C.c_int32), ("graph_used", C.c_int32),
        ("reserved", C.c_int32), ("step_ms", C.c_double), ("total_ms", C.c_double),
    ]


class ServerOpts(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t), ("chunk_steps", C.c_int32),
        ("max_active", C.c_int32), ("precision", C.c_int32), ("policy", C.c_int32),
        ("sparse_density", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
        ("cancelled", C.c_int64), ("active", C.c_int64), ("queued", C.c_int64),
        ("bytes_in_use", C.c_size_t), ("busy_ms", C.c_double), ("chunks", C.c_int64),
        ("max_active_seen", C.c_int64),
    ]


_lib: Optional[C.CDLL] = None


def _bind(lib: C.CDLL) -> None:
    P, I32, U64, D, SZ = C.c_void_p, C.c_int32, C.c_uint64, C.c_double, C.c_size_t
    sig = {
        "sb_version": (C.c_char_p, []),
        "sb_status_string": (C.c_char_p, [C.c_int]),
        "sb_last_error_message": (C.c_char_p, []),
        "sb_default_params": (C.c_int, [C.POINTER(Params)]),
        "sb_edgelist_parse": (C.c_int, [C.c_char_p, SZ, C.c_int, C.POINTER(I32), P, I32]),
        "sb_edgelist_parse_field": (C.c_int, [C.c_char_p, SZ, C.c_int, C.POINTER(I32), P, I32, P,
                                              C.POINTER(I32)]),
        "sb_problem_set_field": (C.c_int, [P, P]),
        "sb_problem_set_sparse": (C.c_int, [P, I32]),
        "sb_problem_sparse_nnz": (C.c_int, [P, C.POINTER(C.c_int64)]),
        "sb_instance_stats": (C.c_int, [I32, P, I32, C.POINTER(D), C.POINTER(I32), C.POINTER(D)]),
        "sb_c0_auto": (D, [I32, I32, D, D]),
        "sb_generate": (C.c_int, [I32, I32, D, U64, P]),
        "sb_ctx_create": (C.c_int, [I32, C.POINTER(P)]),
        "sb_ctx_destroy": (C.c_int, [P]),
        "sb_problem_create": (C.c_int, [P, I32, P, I32, C.POINTER(P)]),
        "sb_problem_from_edgelist_file": (C.c_int, [P, C.c_char_p, C.c_int, C.POINTER(P)]),
        "sb_problem_generate": (C.c_int, [P, I32, I32, D, U64, C.POINTER(P)]),
        "sb_problem_read_j": (C.c_int, [P, P, I32]),
        "sb_problem_destroy": (C.c_int, [P]),
        "sb_problem_info": (C.c_int, [P, C.POINTER(I32), C.POINTER(I32), C.POINTER(D), C.POINTER(I32)]),
        "sb_solve": (C.c_int, [P, P, C.POINTER(Params), C.POINTER(Result)]),
        "sb_step_coeffs_for": (C.c_int, [C.POINTER(Params), D, I32, C.POINTER(StepCoeffs)]),
        "sb_step": (C.c_int, [P, I32, P, I32, I32, I32, P, P, P, I32, I32, C.POINTER(StepCoeffs), P]),
        "sb_field": (C.c_int, [P, P, I32, I32, P, I32, P]),
        "sb_init_state": (C.c_int, [P, I32, U64, I32, I32, I32, P, P, P]),
        "sb_asb_prologue": (C.c_int, [P, C.POINTER(Params), I32, I32, I32, P, P, P]),
        "sb_energy": (C.c_int, [P, P, I32, I32, P, I32, I32, P, P]),
        "sb_energy_noalloc": (C.c_int, [P, P, I32, I32, P, I32, I32, P, P]),
        "sb_batch_solve": (C.c_int, [P, I32, P, C.POINTER(Params), C.POINTER(Result), C.POINTER(BatchStats)]),
        "sb_server_create": (C.c_int, [P, C.POINTER(ServerOpts), C.POINTER(P)]),
        "sb_server_submit": (C.c_int, [P, I32, P, I32, C.POINTER(Params), C.POINTER(U64)]),
        "sb_server_submit_generated": (C.c_int, [P, I32, I32, D, U64, C.POINTER(Params), C.POINTER(U64)]),
        "sb_server_wait": (C.c_int, [P, U64, C.POINTER(Result), D]),
        "sb_server_cancel": (C.c_int, [P, U64]),
        "sb_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
        "sb_server_destroy": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libsbamd.so (raises OSError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"{path} not built: run __graft_entry__.build() or make -C annealing_amd/csrc")
        lib = C.CDLL(path)
        _bind(lib)
        _lib = lib
    return _lib


def _check(rc: int, where: str) -> None:
    if rc != 0:
        detail = _lib.sb_last_error_message().decode() if _lib is not None else ""
        raise SBError(rc, where, detail)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---- host-only helpers ------------------------------------------------------

def parse_edgelist(text: str | bytes, one_based: bool = False) -> np.ndarray:
    """Dense FP32 J from an edge list (first integer N, then i j w triples)."""
    lib = load()
    b = text.encode() if isinstance(text, str) else text
    n = C.c_int32(0)
    _check(lib.sb_edgelist_parse(b, len(b), int(one_based), C.byref(n), None, 0), "sb_edgelist_parse")
    J = np.zeros((n.value, n.value), dtype=np.float32)
    _check(lib.sb_edgelist_parse(b, len(b), int(one_based), C.byref(n), _ptr(J), n.value), "sb_edgelist_parse")
    return J


def parse_edgelist_field(text: str | bytes, one_based: bool = False):
    """(J, h) from an edge list whose "i i w" lines are an external field h_i;
    h is None when the text has none."""
    lib = load()
    b = text.encode() if isinstance(text, str) else text
    n, has = C.c_int32(0), C.c_int32(0)
    _check(lib.sb_edgelist_parse(b, len(b), int(one_based), C.byref(n), None, 0), "sb_edgelist_parse")
    J = np.zeros((n.value, n.value), dtype=np.float32)
    h = np.zeros(n.value, dtype=np.float32)
    _check(lib.sb_edgelist_parse_field(b, len(b), int(one_based), C.byref(n), _ptr(J), n.value, _ptr(h),
                                       C.byref(has)), "sb_edgelist_parse_field")
    return J, (h if has.value else None)


def instance_stats(J: np.ndarray):
    lib = load()
    J = np.ascontiguousarray(J, dtype=np.float32)
    s, integer, mx = C.c_double(), C.c_int32(), C.c_double()
    _check(lib.sb_instance_stats(J.shape[0], _ptr(J), J.shape[1], C.byref(s), C.byref(integer), C.byref(mx)),
           "sb_instance_stats")
    return s.value, bool(integer.value), mx.value


def c0_auto(variant: int, N: int, a0: float, sigma: float) -> float:
    return load().sb_c0_auto(variant, N, a0, sigma)


def generate(kind: int, N: int, p: float = 1.0, seed: int = 0) -> np.ndarray:
    lib = load()
    J = np.empty((N, N), dtype=np.float32)
    _check(lib.sb_generate(kind, N, float(p), int(seed), _ptr(J)), "sb_generate")
    return J


def default_params() -> Params:
    p = Params()
    _check(load().sb_default_params(C.byref(p)), "sb_default_params")
    return p


def make_params(**kw) -> Params:
    p = default_params()
    for k, v in kw.items():
        if k == "variant" and isinstance(v, str):
            v = VARIANTS[v]
        elif k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        elif k == "init" and isinstance(v, str):
            v = INITS[v]
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def step_coeffs(params: Params, c0: float, k: int) -> StepCoeffs:
    c = StepCoeffs()
    _check(load().sb_step_coeffs_for(C.byref(params), c0, k, C.byref(c)), "sb_step_coeffs_for")
    return c


# ---- GPU objects -------------------------------------------------------------

@dataclass
class Solution:
    energy: np.ndarray          # float64 [R]
    energy_i64: Optional[np.ndarray]
    spins: Optional[np.ndarray]  # int8 [R, N]
    x: Optional[np.ndarray]
    y: Optional[np.ndarray]
    best: int
    exact: bool
    c0: float
    step_ms: float
    total_ms: float
    cut: Optional[np.ndarray] = None  # float64 [R], max-cut value of W = -J


def _result_struct(R: int, N: int, spins: bool, state: bool):
    H = np.zeros(R, dtype=np.float64)
    Hi = np.zeros(R, dtype=np.int64)
    S = np.zeros((R, N), dtype=np.int8) if spins else None
    X = np.zeros((R, N), dtype=np.float32) if state else None
    Y = np.zeros((R, N), dtype=np.float32) if state else None
    cut = np.zeros(R, dtype=np.float64)
    res = Result()
    res.energy_f64 = H.ctypes.data_as(C.POINTER(C.c_double))
    res.cut = cut.ctypes.data_as(C.POINTER(C.c_double))
    res.energy_i64 = Hi.ctypes.data_as(C.POINTER(C.c_int64))
    if S is not None:
        res.spins = S.ctypes.data_as(C.POINTER(C.c_int8))
    if X is not None:
        res.x_final = X.ctypes.data_as(C.POINTER(C.c_float))
        res.y_final = Y.ctypes.data_as(C.POINTER(C.c_float))
    return res, (H, Hi, S, X, Y, cut)


def _solution(res: Result, bufs) -> Solution:
    H, Hi, S, X, Y, cut = bufs
    return Solution(H, Hi if res.energy_exact else None, S, X, Y, int(res.best_replica),
                    bool(res.energy_exact), float(res.c0_used), float(res.step_ms), float(res.total_ms), cut)


class Problem:
    def __init__(self, ctx: "Context", handle: C.c_void_p):
        self.ctx = ctx
        self.handle = handle
        ctx._problems.add(self)
        N, integer, sigma, jf = C.c_int32(), C.c_int32(), C.c_double(), C.c_int32()
        _check(load().sb_problem_info(handle, C.byref(N), C.byref(integer), C.byref(sigma), C.byref(jf)),
               "sb_problem_info")
        self.N = N.value
        self.integer = bool(integer.value)
        self.sigma = sigma.value
        self.split_jfmt = jf.value

    def set_field(self, h: np.ndarray | None) -> None:
        if h is None:
            _check(load().sb_problem_set_field(self.handle, None), "sb_problem_set_field")
            return
        h = np.ascontiguousarray(h, dtype=np.float32)
        assert h.shape == (self.N,)
        _check(load().sb_problem_set_field(self.handle, _ptr(h)), "sb_problem_set_field")

    def set_sparse(self, enable: bool = True) -> None:
        """Build (or free) the CSR form; solves then run the sparse family."""
        _check(load().sb_problem_set_sparse(self.handle, int(bool(enable))), "sb_problem_set_sparse")

    @property
    def nnz(self) -> int:
        n = C.c_int64()
        _check(load().sb_problem_sparse_nnz(self.handle, C.byref(n)), "sb_problem_sparse_nnz")
        return n.value

    def read_j(self) -> np.ndarray:
        J = np.empty((self.N, self.N), dtype=np.float32)
        _check(load().sb_problem_read_j(self.handle, _ptr(J), self.N), "sb_problem_read_j")
        return J

    def close(self) -> None:
        # A problem must be destroyed before its context.
        if self.handle and self.ctx.handle:
            load().sb_problem_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One GPU context (stream, workspace, autotune cache)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.sb_ctx_create(device, C.byref(h)), "sb_ctx_create")
        self.handle = h
        self._problems = weakref.WeakSet()

    def close(self) -> None:
        if self.handle:
            for p in list(self._problems):
                p.close()
            self.lib.sb_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def problem(self, J: np.ndarray, h: np.ndarray | None = None) -> Problem:
        """Upload J (and an optional external field h, N floats)."""
        J = np.ascontiguousarray(J, dtype=np.float32)
        ph = C.c_void_p()
        _check(self.lib.sb_problem_create(self.handle, J.shape[0], _ptr(J), J.shape[1], C.byref(ph)),
               "sb_problem_create")
        prob = Problem(self, ph)
        if h is not None:
            prob.set_field(h)
        return prob

    def generate(self, kind: int, N: int, p: float = 1.0, seed: int = 0) -> Problem:
        """Synthetic instance generated on the device (see sb_generate)."""
        h = C.c_void_p()
        _check(self.lib.sb_problem_generate(self.handle, kind, N, float(p), int(seed), C.byref(h)),
               "sb_problem_generate")
        return Problem(self, h)

    def problem_from_file(self, path: str, one_based: bool = False) -> Problem:
        h = C.c_void_p()
        _check(self.lib.sb_problem_from_edgelist_file(self.handle, path.encode(), int(one_based), C.byref(h)),
               "sb_problem_from_edgelist_file")
        return Problem(self, h)

    def solve(self, prob: Problem, spins: bool = True, state: bool = False, **kw) -> Solution:
        params = make_params(**kw)
        res, bufs = _result_struct(params.replicas, prob.N, spins, state)
        _check(self.lib.sb_solve(self.handle, prob.handle, C.byref(params), C.byref(res)), "sb_solve")
        return _solution(res, bufs)

    def field(self, prob: Problem, X: np.ndarray, precision="fp32", variant="bsb") -> np.ndarray:
        """F = J.X (or J.sign(X) for dSB) through the product kernels; X is [R, N]."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        R = X.shape[0]
        F = np.zeros((R, prob.N), dtype=np.float32)
        prec = PRECISIONS[precision] if isinstance(precision, str) else precision
        var = VARIANTS[variant] if isinstance(variant, str) else variant
        _check(self.lib.sb_field(self.handle, prob.handle, prec, var, _ptr(X), R, _ptr(F)), "sb_field")
        return F

    def batch_solve(self, probs, params_list, spins: bool = True, state: bool = False):
        n = len(probs)
        arr_p = (C.c_void_p * n)(*[p.handle for p in probs])
        arr_par = (Params * n)(*params_list)
        arr_res = (Result * n)()
        keep = []
        for i in range(n):
            r, bufs = _result_struct(params_list[i].replicas, probs[i].N, spins, state)
            arr_res[i] = r
            keep.append(bufs)
        st = BatchStats()
        _check(self.lib.sb_batch_solve(self.handle, n, arr_p, arr_par, arr_res, C.byref(st)), "sb_batch_solve")
        return [_solution(arr_res[i], keep[i]) for i in range(n)], st


class Server:
    """Continuous-batching request server (sb_server_* of include/sb.h)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, chunk_steps: int = 50, max_active: int = 0,
                 precision="split", policy="shortest", sparse_density: float = 0.0):
        self.ctx = ctx
        self.lib = ctx.lib
        o = ServerOpts()
        o.hbm_budget_bytes = int(hbm_budget_bytes)
        o.chunk_steps = chunk_steps
        o.max_active = max_active
        o.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
        o.policy = SERVE_POLICIES[policy] if isinstance(policy, str) else policy
        o.sparse_density = float(sparse_density)
        h = C.c_void_p()
        _check(self.lib.sb_server_create(ctx.handle, C.byref(o), C.byref(h)), "sb_server_create")
        self.handle = h
        self._shape = {}

    def submit(self, J: np.ndarray, **kw) -> int:
        J = np.ascontiguousarray(J, dtype=np.float32)
        p = make_params(**kw)
        t = C.c_uint64()
        _check(self.lib.sb_server_submit(self.handle, J.shape[0], _ptr(J), J.shape[1], C.byref(p), C.byref(t)),
               "sb_server_submit")
        self._shape[t.value] = (p.replicas, J.shape[0])
        return t.value

    def submit_generated(self, kind: int, N: int, p: float, gen_seed: int, **kw) -> int:
        """Instance (kind, N, p, gen_seed) generated on the device; kw are solver params."""
        par = make_params(**kw)
        t = C.c_uint64()
        _check(self.lib.sb_server_submit_generated(self.handle, kind, N, float(p), int(gen_seed), C.byref(par),
                                                   C.byref(t)), "sb_server_submit_generated")
        self._shape[t.value] = (par.replicas, N)
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0, spins: bool = True) -> Solution:
        R, N = self._shape.get(ticket, (1, 1))
        res, bufs = _result_struct(R, N, spins, False)
        try:
            _check(self.lib.sb_server_wait(self.handle, ticket, C.byref(res), float(timeout_ms)), "sb_server_wait")
        except SBError as e:
            if e.status != 10:  # the ticket stays alive only after a timeout
                self._shape.pop(ticket, None)
            raise
        self._shape.pop(ticket, None)
        return _solution(res, bufs)

    def cancel(self, ticket: int) -> None:
        _check(self.lib.sb_server_cancel(self.handle, ticket), "sb_server_cancel")

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(self.lib.sb_server_stats_get(self.handle, C.byref(st)), "sb_server_stats_get")
        return st

    def close(self) -> None:
        if self.handle:
            self.lib.sb_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
