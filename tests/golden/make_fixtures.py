"""Extract the reference's own golden data into small, pickle-free fixtures.

Sources (read as data only; nothing from the reference is imported or run):
  tests/test_julia/LANCZOS_STEP10.npz      save_status of a 10-step
      biconjugate (use_wigner=False) Lanczos on the 16-atom SnTe ensemble,
      full 384-operation symmetry (written by tests/test_lanczos_fast/
      test_julia_lf.py:24-45).  Object-valued keys (sym_julia, deg_julia)
      are skipped because they need pickle.
  tests/test_julia/test_c_code_lanczos/mode10.{ndegs,blockid,block0-8}
      the degenerate-block layout of the same system.
  tests/test_julia/data/xats_pop1.npy      atomic positions (Angstrom) of
      the 10 configurations, used to recover the supercell atom order.
  tests/test_julia/data/dyn_gen_pop1_1     header: alat and unit cell.

  tests/test_julia/data/dyn_gen_pop1_{1,2,3}, xats_pop1.npy,
      forces_pop1.npy   the raw inputs of that run (the dynamical-matrix
      files of the 3 irreducible q-points and the ensemble in the
      load_bin layout), copied unchanged into tests/golden/snte_raw/ for the
      file readers and the preprocessing.

  tests/test_julia/test_c_code_lanczos/mode10.{json,freqs,masses,rho,ndegs,
      blockid,block0-8,psi}   the input files the reference's
      prepare_input_files wrote for the C++ driver on the same system
      (mode 10, 5 steps), copied unchanged into tests/golden/snte_c_input/.

Run:  python tests/golden/make_fixtures.py /root/reference
"""
import os
import sys

import numpy as np


def main(ref_root):
    tj = os.path.join(ref_root, "tests", "test_julia")
    out = {}
    with np.load(os.path.join(tj, "LANCZOS_STEP10.npz"), allow_pickle=False) as d:
        for k in d.files:
            try:
                out[k] = d[k]
            except ValueError:
                pass  # object array: needs pickle, not loaded
    cdir = os.path.join(tj, "test_c_code_lanczos")
    out["ref_ndegs"] = np.loadtxt(os.path.join(cdir, "mode10.ndegs"), dtype=np.int32)
    out["ref_blockid"] = np.loadtxt(os.path.join(cdir, "mode10.blockid"), dtype=np.int32)
    blocks = [np.atleast_1d(np.loadtxt(os.path.join(cdir, "mode10.block%d" % i), dtype=np.int32))
              for i in range(len(out["ref_ndegs"]))]
    out["ref_blocks_concat"] = np.concatenate(blocks)
    out["xats"] = np.load(os.path.join(tj, "data", "xats_pop1.npy"), allow_pickle=False)
    with open(os.path.join(tj, "data", "dyn_gen_pop1_1")) as fp:
        lines = fp.read().splitlines()
    alat = float(lines[2].split()[3])
    cell = np.array([[float(x) for x in lines[4 + i].split()] for i in range(3)])
    out["alat_bohr"] = np.float64(alat)
    out["unit_cell_alat"] = cell
    # atomic positions in alat units (two atoms) from the header
    pos = []
    for ln in lines[9:11]:
        parts = ln.split()
        pos.append([float(x) for x in parts[2:5]])
    out["unit_pos_alat"] = np.array(pos)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "snte_step10.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, sorted(out))
    raw = os.path.join(os.path.dirname(os.path.abspath(__file__)), "snte_raw")
    os.makedirs(raw, exist_ok=True)
    for i in (1, 2, 3):
        name = "dyn_gen_pop1_%d" % i
        with open(os.path.join(tj, "data", name)) as fp, open(os.path.join(raw, name), "w") as fo:
            fo.write(fp.read())
    for name in ("xats_pop1.npy", "forces_pop1.npy"):
        np.save(os.path.join(raw, name), np.load(os.path.join(tj, "data", name), allow_pickle=False))
    print("wrote", raw, sorted(os.listdir(raw)))
    cin = os.path.join(os.path.dirname(os.path.abspath(__file__)), "snte_c_input")
    os.makedirs(cin, exist_ok=True)
    names = ["mode10.%s" % e for e in ("json", "freqs", "masses", "rho", "ndegs", "blockid", "psi")]
    names += ["mode10.block%d" % i for i in range(len(out["ref_ndegs"]))]
    for name in names:
        with open(os.path.join(cdir, name)) as fp, open(os.path.join(cin, name), "w") as fo:
            fo.write(fp.read())
    print("wrote", cin, sorted(os.listdir(cin)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
