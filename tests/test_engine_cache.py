"""The drop-in binding's ensemble cache is keyed on content (CPU, no device).

GetPerturbAverageSym re-uses an uploaded ensemble only when X, Y, w, rho, the
symmetry blocks, the degenerate spaces and the block ids hold the same
bytes; an in-place edit of any of them, or other symmetry matrices of the
same shape, must give another key.
"""
import numpy as np

from tdscha_amd.engine import _digest
from tests.systems import random_block_system


def _key(s, rho=None, syms=None):
    return _digest([s["X"], s["Y"], s["w"], s["rho"] if rho is None else rho,
                    np.asarray(s["sbi"], np.int32)]
                   + list(s["syms"] if syms is None else syms)
                   + [np.asarray(b, np.int32) for b in s["blocks"]])


def test_key_changes_with_content_not_address():
    s = random_block_system(20, 50, 3, [2, 2] + [1] * 16, seed=1)
    k0 = _key(s)
    assert _key(s) == k0
    # a copy at another address has the same key
    assert _key({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()}) == k0
    rho = s["rho"]
    rho[3] *= 2.0                       # in place
    k1 = _key(s)
    assert k1 != k0
    s["X"][0, 0] += 1e-12               # in place, tiny
    assert _key(s) != k1
    s2 = random_block_system(20, 50, 3, [2, 2] + [1] * 16, seed=2)
    assert _key(s, syms=s2["syms"]) != _key(s)   # same shapes, other matrices
    blocks = [b.copy() for b in s["blocks"]]
    blocks[0] = blocks[0][::-1].copy()
    assert _digest([np.asarray(b, np.int32) for b in blocks]) != \
        _digest([np.asarray(b, np.int32) for b in s["blocks"]])
