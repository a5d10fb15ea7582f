"""Test systems: the reference's SnTe golden ensemble and seeded random ones."""
import os

import numpy as np

from tdscha_amd import symmetry as SY

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOHR_PER_ANGSTROM = 1.889725989


def load_snte():
    """SnTe 2x2x2 (16 atoms, 45 modes, 10 configs, T=250 K) + its 384 symmetries."""
    d = dict(np.load(os.path.join(GOLDEN, "snte_step10.npz")))
    alat = float(d["alat_bohr"])
    cell = d["unit_cell_alat"] * alat
    upos = d["unit_pos_alat"] * alat
    lat = 2 * cell
    xat = d["xats"][0] * BOHR_PER_ANGSTROM
    sites = np.array([upos[a] + i * cell[0] + j * cell[1] + k * cell[2]
                      for i in range(2) for j in range(2) for k in range(2) for a in range(2)])
    inv = np.linalg.inv(lat)
    pos = np.zeros_like(xat)
    for ia, x in enumerate(xat):
        dd = (sites - x) @ inv
        dd -= np.round(dd)
        dist = np.linalg.norm(dd @ lat, axis=1)
        j = int(np.argmin(dist))
        pos[ia] = x + dd[j] @ lat
    trans = [i * cell[0] + j * cell[1] + k * cell[2]
             for i in range(2) for j in range(2) for k in range(2)]
    d["positions"] = pos
    d["lattice"] = lat
    d["translations"] = trans
    d["syms_full"] = SY.build_mode_symmetries(pos, lat, pos[0], d["pols"], d["w"], trans,
                                              gamma_only=False)
    d["syms_gamma"] = SY.build_mode_symmetries(pos, lat, pos[0], d["pols"], d["w"], trans,
                                               gamma_only=True)
    return d


def random_block_system(n, N, n_syms, block_sizes, seed=0, T=300.0):
    """Seeded ensemble with random orthogonal blocks (not a group: kernel parity only)."""
    rng = np.random.default_rng(seed)
    assert sum(block_sizes) == n
    w = np.sort(rng.uniform(2e-4, 2e-3, size=n))
    # make blocks degenerate
    start = 0
    blocks = []
    for d in block_sizes:
        w[start:start + d] = w[start]
        blocks.append(np.arange(start, start + d, dtype=np.int32))
        start += d
    X = rng.normal(size=(N, n)) * 20.0
    Y = rng.normal(size=(N, n)) * 1e-6
    rho = rng.uniform(0.5, 1.5, size=N)
    syms = []
    for d in block_sizes:
        s = np.zeros((n_syms, d, d))
        s[0] = np.eye(d)
        for j in range(1, n_syms):
            q, _ = np.linalg.qr(rng.normal(size=(d, d)))
            s[j] = q
        syms.append(s)
    sbi = np.zeros(n, dtype=np.int32)
    for b, modes in enumerate(blocks):
        sbi[modes] = b
    return dict(X=X, Y=Y, w=w, rho=rho, T=T, syms=syms, blocks=blocks, sbi=sbi)


def random_perturbation(n, seed=1, scale=1.0):
    rng = np.random.default_rng(seed)
    R1 = rng.normal(size=n) * scale
    Y1 = rng.normal(size=(n, n)) * scale
    return R1, Y1 + Y1.T
