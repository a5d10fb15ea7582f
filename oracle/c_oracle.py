"""ctypes access to the two C oracles.  TEST INFRASTRUCTURE ONLY.

* ``libtdscha_oracle.so``  - the project's C restatement (oracle/perturb_avg.c)
* ``_ref/libref_lanczos.so`` - the reference's own LanczosFunctions.c compiled
  from its sources by ``make -C oracle ref`` (only where /root/reference
  exists; the built .so travels with the tree).

``ref_perturb_averages`` calls the three reference kernels in the order and
with the overwrite/accumulate semantics of the CPython binding
Get_Perturb_AveragesSym (CModules/odd_corr_module.c:491-634): f overwritten,
d2v overwritten by the R1 kernel, then the D4 kernel adds to it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_P = np.ctypeslib.ndpointer


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


_lib = None
_ref = None


def oracle_lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libtdscha_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle C library missing: run `make -C oracle`")
        _lib = C.CDLL(path)
        _lib.oracle_perturb_averages.restype = C.c_int
    return _lib


def ref_lib():
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libref_lanczos.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
        for name in ("get_f_average_from_Y_pert_sym_fast",
                     "get_d2v_dR2_from_R_pert_sym_fast",
                     "get_d2v_dR2_from_Y_pert_sym_fast"):
            getattr(_ref, name).restype = None
    return _ref


def _flat_blocks(block_syms, block_modes):
    dims = np.array([len(b) for b in block_modes], dtype=np.int32)
    modes = np.ascontiguousarray(np.concatenate([np.asarray(b, np.int32) for b in block_modes]))
    syms = np.ascontiguousarray(np.concatenate([np.asarray(s, np.float64).ravel() for s in block_syms]))
    return dims, modes, syms


def oracle_perturb_averages(X, Y, w, rho, R1, Y1, T, apply_d4, block_syms,
                            block_modes, sym_block_id, b_begin=0, b_end=-1,
                            n_threads=1):
    """C restatement; same semantics as lanczos_oracle.perturb_averages."""
    lib = oracle_lib()
    X = np.ascontiguousarray(X, np.float64)
    Y = np.ascontiguousarray(Y, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    rho = np.ascontiguousarray(rho, np.float64)
    R1 = np.ascontiguousarray(R1, np.float64)
    Y1 = np.ascontiguousarray(Y1, np.float64)
    sbi = np.ascontiguousarray(sym_block_id, np.int32)
    dims, modes, syms = _flat_blocks(block_syms, block_modes)
    N, n = X.shape
    n_syms = np.asarray(block_syms[0]).shape[0]
    f = np.zeros(n)
    d2v = np.zeros((n, n))
    rc = lib.oracle_perturb_averages(
        C.c_int(n), C.c_int(N), _dp(X), _dp(Y), _dp(w), _dp(rho), _dp(R1), _dp(Y1),
        C.c_double(T), C.c_int(int(apply_d4)), C.c_int(n_syms), C.c_int(len(dims)),
        _ip(dims), _ip(modes), _dp(syms), _ip(sbi), C.c_long(b_begin), C.c_long(b_end),
        C.c_int(n_threads), _dp(f), _dp(d2v))
    if rc != 0:
        raise ValueError("oracle_perturb_averages failed with code %d" % rc)
    return f, d2v


class _Quiet:
    """Silence the reference kernel's per-call printf on fd 1."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *a):
        C.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


def ref_perturb_averages(X, Y, w, rho, R1, Y1, T, apply_d4, block_syms,
                         block_modes, sym_block_id):
    """The reference C kernels themselves (None if oracle/_ref is not built)."""
    lib = ref_lib()
    if lib is None:
        return None
    X = np.ascontiguousarray(X, np.float64)
    Y = np.ascontiguousarray(Y, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    rho = np.ascontiguousarray(rho, np.float64)
    R1 = np.ascontiguousarray(R1, np.float64)
    Y1 = np.ascontiguousarray(Y1, np.float64)
    sbi = np.ascontiguousarray(sym_block_id, np.int32)
    N, n = X.shape
    syms = [np.ascontiguousarray(s, np.float64) for s in block_syms]
    modes = [np.ascontiguousarray(b, np.int32) for b in block_modes]
    n_syms = syms[0].shape[0]
    ndeg = np.array([len(b) for b in modes], dtype=np.int32)
    sym_ptrs = (C.POINTER(C.c_double) * len(syms))(*[_dp(s) for s in syms])
    deg_ptrs = (C.POINTER(C.c_int) * len(modes))(*[_ip(m) for m in modes])
    f = np.zeros(n)
    d2v = np.zeros((n, n))
    args_common = (C.c_double(T), C.c_int(n), C.c_int(N), _dp(rho), sym_ptrs,
                   C.c_int(n_syms), _ip(ndeg), deg_ptrs, _ip(sbi))
    with _Quiet():
        lib.get_f_average_from_Y_pert_sym_fast(_dp(X), _dp(Y), _dp(w), _dp(Y1), *args_common, _dp(f))
        lib.get_d2v_dR2_from_R_pert_sym_fast(_dp(X), _dp(Y), _dp(w), _dp(R1), *args_common, _dp(d2v))
        if apply_d4:
            lib.get_d2v_dR2_from_Y_pert_sym_fast(_dp(X), _dp(Y), _dp(w), _dp(Y1), *args_common, _dp(d2v))
    return f, d2v
