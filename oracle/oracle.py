"""ctypes binding of oracle/liboracle.so (the CPU restatement in sb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SB_ORACLE_LIB selects another build of the same source (the sanitized one of
# `make -C oracle sanitize`).
LIB_PATH = os.environ.get("SB_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

ASB, BSB, DSB, ASB_M1 = 0, 1, 2, 3
ORDER_SEQ, ORDER_LANES, ORDER_MFMA, ORDER_SPLIT = 0, 1, 2, 3  # field orders (sb_oracle.h)

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built: make -C oracle")
        lib = C.CDLL(LIB_PATH)
        P, I, U32, U64, D, F = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_double, C.c_float
        sig = {
            "orc_xorshift32": (U32, [C.POINTER(U32)]),
            "orc_ref_draw_value": (F, [U32]),
            "orc_init_reference": (None, [I, I, I, U32, P, P]),
            "orc_keyed_u24": (U32, [U64, U32, U32, U32]),
            "orc_init_uniform": (None, [I, I, I, U64, P, P]),
            "orc_sigma": (D, [I, P, I]),
            "orc_c0_auto": (D, [I, I, D, P, I]),
            "orc_coeffs": (None, [I, I, I, D, D, D, P]),
            "orc_field_f32": (None, [I, I, P, I, P, I, I, P]),
            "orc_field_f64": (None, [I, I, P, I, P, I, I, P]),
            "orc_field_f32_lanes": (None, [I, I, P, I, P, I, I, P]),
            "orc_field_f32_mfma": (None, [I, I, P, I, P, I, I, P]),
            "orc_field_split": (None, [I, I, P, I, P, I, I, P]),
            "orc_mfma_bf16": (F, [F, P, P]),
            "orc_bsb_update_f32": (None, [F, F, F, F, C.POINTER(F), C.POINTER(F)]),
            "orc_asb_local_f32": (None, [C.POINTER(F), C.POINTER(F), F, D, D, D, I]),
            "orc_m1_x_f32": (F, [F, F, D, D]),
            "orc_m1_y_f32": (F, [F, F, F, F, D, D, D, D]),
            "orc_run_f32": (I, [I, I, I, P, I, P, P, I, I, I, I, D, D, D, D, I, I]),
            "orc_run_f64": (I, [I, I, I, P, I, P, P, I, I, I, I, D, D, D, D, I, I]),
            "orc_run_f32_order": (I, [I, I, I, I, P, I, P, P, P, I, I, I, I, D, D, D, D, I, I]),
            "orc_run_f64_h": (I, [I, I, I, P, I, P, P, P, I, I, I, I, D, D, D, D, I, I]),
            "orc_energy_field_i64": (I, [I, I, P, P, P]),
            "orc_energy_field_f64": (None, [I, I, P, P, P]),
            "orc_edgelist_fill_h": (I, [C.c_char_p, C.c_size_t, I, I, P, I, P]),
            "orc_spins_f32": (None, [I, I, P, I, P]),
            "orc_energy_i64": (I, [I, I, P, I, P, P]),
            "orc_energy_f64": (None, [I, I, P, I, P, P]),
            "orc_ground_state": (I, [I, P, I, C.POINTER(D), P]),
            "orc_edgelist_header": (I, [C.c_char_p, C.c_size_t, C.POINTER(I)]),
            "orc_edgelist_fill": (I, [C.c_char_p, C.c_size_t, I, I, P, I]),
            "orc_has_fma": (I, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def pad_to(N: int, m: int = 64) -> int:
    return (N + m - 1) // m * m


def xorshift32(state: int) -> int:
    s = C.c_uint32(state)
    load().orc_xorshift32(C.byref(s))
    return s.value


def ref_draw_value(v: int) -> float:
    return load().orc_ref_draw_value(v)


def init_state(kind: str, N: int, R: int, seed: int = 0, ld: int | None = None):
    """(x, y) float32 [R, ld] from the reference (xorshift) or keyed uniform init."""
    ld = ld or pad_to(N)
    x = np.zeros((R, ld), np.float32)
    y = np.zeros((R, ld), np.float32)
    if kind == "reference":
        load().orc_init_reference(N, R, ld, seed & 0xFFFFFFFF, _p(x), _p(y))
    else:
        load().orc_init_uniform(N, R, ld, seed, _p(x), _p(y))
    return x, y


def keyed_u24(seed, replica, spin, comp):
    return load().orc_keyed_u24(seed, replica, spin, comp)


def sigma(J):
    J = np.ascontiguousarray(J, np.float32)
    return load().orc_sigma(J.shape[0], _p(J), J.shape[1])


def c0_auto(variant, J, a0=1.0):
    J = np.ascontiguousarray(J, np.float32)
    return load().orc_c0_auto(variant, J.shape[0], a0, _p(J), J.shape[1])


def coeffs(variant, k, steps, dt, a0, c0):
    out = np.zeros(4, np.float32)
    load().orc_coeffs(variant, k, steps, dt, a0, c0, _p(out))
    return out


def field_f32(J, x, use_sign=False, order=ORDER_SEQ):
    J = np.ascontiguousarray(J, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    N = J.shape[0]
    R, ld = x.shape
    F = np.zeros((R, ld), np.float32)
    fn = {ORDER_SEQ: load().orc_field_f32, ORDER_LANES: load().orc_field_f32_lanes,
          ORDER_MFMA: load().orc_field_f32_mfma, ORDER_SPLIT: load().orc_field_split}[order]
    fn(N, R, _p(J), J.shape[1], _p(x), ld, int(use_sign), _p(F))
    return F


def mfma_bf16(c, a, b):
    """The oracle's model of one v_mfma_f32_32x32x16_bf16 output: c plus the
    16 products of BF16 values a[k] b[k] (see orc_mfma_bf16)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == (16,) and b.shape == (16,)
    return np.float32(load().orc_mfma_bf16(float(c), _p(a), _p(b)))


def field_f64(J, x, use_sign=False):
    J = np.ascontiguousarray(J, np.float32)
    x = np.ascontiguousarray(x, np.float64)
    N = J.shape[0]
    R, ld = x.shape
    F = np.zeros((R, ld), np.float64)
    load().orc_field_f64(N, R, _p(J), J.shape[1], _p(x), ld, int(use_sign), _p(F))
    return F


def bsb_update_f32(F, alpha, beta, gamma, x, y):
    xx, yy = C.c_float(x), C.c_float(y)
    load().orc_bsb_update_f32(F, alpha, beta, gamma, C.byref(xx), C.byref(yy))
    return xx.value, yy.value


def asb_local_f32(x, y, p, a0, K, dts, M):
    xx, yy = C.c_float(x), C.c_float(y)
    load().orc_asb_local_f32(C.byref(xx), C.byref(yy), p, a0, K, dts, M)
    return xx.value, yy.value


def m1_x_f32(x, y, a0, dt):
    return load().orc_m1_x_f32(x, y, a0, dt)


def m1_y_f32(F, x, y, p, a0, K, dt, xi):
    return load().orc_m1_y_f32(F, x, y, p, a0, K, dt, xi)


def _hp(h, N):
    if h is None:
        return None, None
    h = np.ascontiguousarray(h, np.float32)
    assert h.shape == (N,)
    return h, _p(h)


def run_f32(variant, J, x, y, k0, k1, steps, dt=1.0, a0=1.0, c0=0.0, K=1.0, M=2, nthreads=1,
            order=ORDER_SEQ, h=None):
    """Advance float32 state (x, y) [R, ld] in place over steps [k0, k1)."""
    J = np.ascontiguousarray(J, np.float32)
    assert x.dtype == np.float32 and y.dtype == np.float32 and x.flags.c_contiguous
    R, ld = x.shape
    hk, hptr = _hp(h, J.shape[0])
    rc = load().orc_run_f32_order(order, variant, J.shape[0], R, _p(J), J.shape[1], hptr, _p(x), _p(y), ld,
                                  k0, k1, steps, dt, a0, c0, K, M, nthreads)
    if rc:
        raise RuntimeError(f"orc_run_f32 failed: {rc}")


def run_f64(variant, J, x, y, k0, k1, steps, dt=1.0, a0=1.0, c0=0.0, K=1.0, M=2, nthreads=1, h=None):
    J = np.ascontiguousarray(J, np.float32)
    assert x.dtype == np.float64 and y.dtype == np.float64 and x.flags.c_contiguous
    R, ld = x.shape
    hk, hptr = _hp(h, J.shape[0])
    rc = load().orc_run_f64_h(variant, J.shape[0], R, _p(J), J.shape[1], hptr, _p(x), _p(y), ld, k0, k1,
                              steps, dt, a0, c0, K, M, nthreads)
    if rc:
        raise RuntimeError(f"orc_run_f64 failed: {rc}")


def spins(x, N):
    x = np.ascontiguousarray(x, np.float32)
    R, ld = x.shape
    s = np.zeros((R, N), np.int8)
    load().orc_spins_f32(N, R, _p(x), ld, _p(s))
    return s


def energy_i64(J, s, h=None):
    J = np.ascontiguousarray(J, np.float32)
    s = np.ascontiguousarray(s, np.int8)
    R, N = s.shape
    H = np.zeros(R, np.int64)
    rc = load().orc_energy_i64(N, R, _p(J), J.shape[1], _p(s), _p(H))
    if rc:
        raise ValueError("J is not integer-valued")
    if h is not None:
        hk, hptr = _hp(h, N)
        if load().orc_energy_field_i64(N, R, hptr, _p(s), _p(H)):
            raise ValueError("h is not integer-valued")
    return H


def energy_f64(J, s, h=None):
    J = np.ascontiguousarray(J, np.float32)
    s = np.ascontiguousarray(s, np.int8)
    R, N = s.shape
    H = np.zeros(R, np.float64)
    load().orc_energy_f64(N, R, _p(J), J.shape[1], _p(s), _p(H))
    if h is not None:
        hk, hptr = _hp(h, N)
        load().orc_energy_field_f64(N, R, hptr, _p(s), _p(H))
    return H


def ground_state(J):
    J = np.ascontiguousarray(J, np.float32)
    N = J.shape[0]
    H = C.c_double()
    s = np.zeros(N, np.int8)
    rc = load().orc_ground_state(N, _p(J), J.shape[1], C.byref(H), _p(s))
    if rc:
        raise ValueError("N too large for brute force")
    return H.value, s


def parse_edgelist(text, one_based=False):
    b = text.encode() if isinstance(text, str) else text
    n = C.c_int()
    rc = load().orc_edgelist_header(b, len(b), C.byref(n))
    if rc:
        return rc, None
    J = np.zeros((n.value, n.value), np.float32)
    rc = load().orc_edgelist_fill(b, len(b), int(one_based), n.value, _p(J), n.value)
    return rc, (J if rc == 0 else None)


def parse_edgelist_field(text, one_based=False):
    """(rc, J, h): "i i w" lines read as the external field h."""
    b = text.encode() if isinstance(text, str) else text
    n = C.c_int()
    rc = load().orc_edgelist_header(b, len(b), C.byref(n))
    if rc:
        return rc, None, None
    J = np.zeros((n.value, n.value), np.float32)
    h = np.zeros(n.value, np.float32)
    rc = load().orc_edgelist_fill_h(b, len(b), int(one_based), n.value, _p(J), n.value, _p(h))
    return rc, (J if rc == 0 else None), (h if rc == 0 else None)


def solve_f32(variant, J, R, steps, init="uniform", seed=0, dt=1.0, a0=1.0, c0=None, K=1.0, M=2,
              nthreads=1, order=ORDER_SEQ, h=None):
    """Whole solve on the CPU: init, (aSB prologue), steps; returns x, y [R, ld]."""
    N = J.shape[0]
    if c0 is None or c0 <= 0:
        c0 = c0_auto(variant, J, a0)
    x, y = init_state(init, N, R, seed)
    run_f32(variant, J, x, y, 0, steps, steps, dt, a0, c0, K, M, nthreads, order, h)
    return x, y, c0
