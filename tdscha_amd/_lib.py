"""ctypes binding of libtdscha_amd.so (the gfx950 HIP library).

There is no CPU fallback: if the library is missing or cannot be loaded,
``lib()`` raises.  Build it with ``python -m tdscha_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtdscha_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "tdscha_amd.h")

_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class TdschaError(RuntimeError):
    pass


def _signatures(L):
    L.tdscha_last_error.restype = C.c_char_p
    L.tdscha_ctx_create.argtypes = [C.POINTER(_vp), C.c_int, C.c_int, _vp, _vp, _vp, _vp, C.c_double,
                                    C.c_int, C.c_int, _vp, _vp, _vp, _vp, C.c_int]
    L.tdscha_ctx_destroy.argtypes = [_vp]
    L.tdscha_ctx_destroy.restype = None
    L.tdscha_ctx_info.argtypes = [_vp, _ip, _ip, C.POINTER(C.c_long), C.POINTER(C.c_long), _ip]
    L.tdscha_synchronize.argtypes = [_vp]
    L.tdscha_perturb_averages.argtypes = [_vp, _vp, _vp, C.c_int, _vp, _vp]
    L.tdscha_perturb_averages_dev.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp]
    L.tdscha_perturb_averages_batched.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, _vp, _vp]
    L.tdscha_perturb_averages_range.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int,
                                                C.c_long, C.c_long, _vp, _vp, _vp]
    L.tdscha_perturb_averages_noalloc.argtypes = [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                                  _vp, _vp, _vp, _vp, C.c_long, C.c_long, _vp,
                                                  _vp, _vp]
    L.tdscha_set_translations.argtypes = [_vp, C.c_int, _vp, C.c_int, _vp, _vp]
    L.tdscha_apply_L.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_int, C.c_double, C.c_int]
    L.tdscha_apply_L_dev.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_double, C.c_int, _vp]
    L.tdscha_apply_L1_inverse.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int]
    L.tdscha_lanczos_run.argtypes = [_vp, C.c_int, _vp, C.c_int, _vp, C.c_double, C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                     _vp, _vp]
    L.tdscha_lanczos_run_biconjugate.argtypes = [_vp, C.c_int, _vp, C.c_int, _vp, C.c_double,
                                                 C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.c_int, C.c_double, C.c_int, _vp, _vp, _vp,
                                                 _vp, _vp]
    L.tdscha_continued_fraction.argtypes = [C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, _vp,
                                            C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                            C.c_int, C.c_double, C.c_double, _vp]
    L.tdscha_profile_perturb_averages.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp]
    L.tdscha_set_ksplit.argtypes = [_vp, C.c_int]
    L.tdscha_autotune.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _ip, C.POINTER(C.c_double)]
    L.tdscha_device_alloc.argtypes = [C.POINTER(_vp), C.c_size_t]
    L.tdscha_device_free.argtypes = [_vp]
    L.tdscha_memcpy.argtypes = [_vp, _vp, C.c_size_t, C.c_int]
    L.tdscha_device_count.argtypes = [_ip]
    L.tdscha_qctx_create.argtypes = [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                     _vp, _vp, C.c_double]
    L.tdscha_qctx_destroy.argtypes = [_vp]
    L.tdscha_qctx_destroy.restype = None
    L.tdscha_qctx_symmetrize.argtypes = [_vp, C.c_int, _vp, _vp]
    L.tdscha_qspace_psi_size.argtypes = [C.c_int, C.c_int, _vp, _ip]
    L.tdscha_qspace_static_psi_size.argtypes = [C.c_int, C.c_int, _vp, _ip]
    L.tdscha_qspace_apply_L_static.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, C.c_int,
                                               C.c_int, C.c_int]
    L.tdscha_qspace_apply_L.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.c_int,
                                        C.c_int]
    L.tdscha_qspace_lanczos_run.argtypes = [_vp, C.c_int, _vp, _vp, _vp, C.c_int, _vp, C.c_double,
                                            C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                            _vp, C.POINTER(C.c_double)]
    L.tdscha_qspace_set_tiles.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.tdscha_qspace_set_d2v_form.argtypes = [_vp, C.c_int]
    L.tdscha_qspace_kpm_run.argtypes = [_vp, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _vp,
                                        _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                        _vp]
    L.tdscha_qspace_get_d2v_form.argtypes = [_vp, _vp]
    L.tdscha_qspace_get_tiles.argtypes = [_vp, _vp]
    L.tdscha_qspace_autotune.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, _vp,
                                         C.POINTER(C.c_double)]
    L.tdscha_mlp_create.argtypes =[C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    _vp]
    L.tdscha_mlp_destroy.argtypes = [_vp]
    L.tdscha_mlp_destroy.restype = None
    L.tdscha_mlp_n_params.argtypes = [_vp, C.POINTER(C.c_long)]
    L.tdscha_mlp_get_params.argtypes = [_vp, _vp]
    L.tdscha_mlp_set_params.argtypes = [_vp, _vp, C.c_int]
    L.tdscha_mlp_forward.argtypes = [_vp, C.c_int, _vp, _vp]
    L.tdscha_mlp_set_mse_weights.argtypes = [_vp, _vp]
    L.tdscha_mlp_set_cf_loss.argtypes = [_vp, C.c_int, C.c_int, _vp, C.c_double, _vp, _vp,
                                         C.c_double]
    L.tdscha_mlp_loss_grad.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp,
                                       C.POINTER(C.c_float), _vp]
    L.tdscha_mlp_adam_step.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, C.c_float,
                                       C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)]
    L.tdscha_mlp_train.argtypes = [_vp, C.c_int, C.c_long, _vp, _vp, _vp, _vp, C.c_int, C.c_int,
                                   C.c_float, C.c_float, C.c_float, C.c_float, C.c_ulonglong, _vp]
    return L


class DeviceBuffer:
    """A device allocation owned through the library (no other GPU runtime)."""

    def __init__(self, shape, dtype="float64"):
        import numpy as np
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().tdscha_device_alloc(C.byref(p), self.nbytes), "tdscha_device_alloc")
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        b = cls(a.shape, a.dtype)
        b.upload(a)
        return b

    def upload(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().tdscha_memcpy(self.ptr, C.c_void_p(a.ctypes.data), self.nbytes, 1), "upload")

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().tdscha_memcpy(C.c_void_p(out.ctypes.data), self.ptr, self.nbytes, 2), "download")
        return out

    def free(self):
        if self.ptr:
            lib().tdscha_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count():
    n = C.c_int(0)
    lib().tdscha_device_count(C.byref(n))
    return n.value


def declared_symbols():
    """Function names declared in include/tdscha_amd.h."""
    with open(HEADER) as fp:
        txt = fp.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tdscha_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TdschaError("libtdscha_amd.so is not built (run `python -m tdscha_amd.build`); "
                              "there is no CPU fallback")
        _lib = _signatures(C.CDLL(LIB_PATH))
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().tdscha_last_error().decode(errors="replace")
        raise TdschaError("%s failed (code %d): %s" % (what, rc, msg))
