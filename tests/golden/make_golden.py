"""Writes the golden fixtures of tests/golden/ from the reference's own host C
(compiled into oracle/_ref/ by oracle/Makefile from the reference tree).

Run in the build container only (the reference tree is not on the GPU box):
    make -C oracle all ref && python tests/golden/make_golden.py

For every case it stores the CSR the reference reader produced (shipped .mtx
files) or the synthetic CSR given to it, and the reference's Tile_create /
blockspmv_cpu / encode outputs for BLOCK_SIZE 16 and 8.  The two small .mtx
data files used by the reader tests are copied next to the fixtures.
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from orc import Ref, REF_ARRAYS  # noqa: E402

GINKGO = "/root/reference/Performance-Comparison/ginkgo-develop/matrices/test"
PETSC = "/root/reference/Performance-Comparison/PETSc/share/petsc/datafiles/matrices"
SHIPPED = {
    "1138_bus": os.path.join(GINKGO, "1138_bus.mtx"),
    "ani1_nonsymm": os.path.join(GINKGO, "ani1_nonsymm.mtx"),
    "isai_spd": os.path.join(GINKGO, "isai_spd.mtx"),
    "nontrivial_mc64_example": os.path.join(GINKGO, "nontrivial_mc64_example.mtx"),
    "LFAT5": os.path.join(PETSC, "LFAT5.mtx"),
}
COPY_MTX = ("1138_bus", "ani1_nonsymm", "LFAT5")


def synthetic_cases():
    """CSR inputs that exercise the layout's corner cases."""
    rng = np.random.default_rng(7)
    cases = {}
    # uint8 wrap: duplicates put 300 entries into one 16x16 tile (row 0 alone has 280)
    rows = [[(c % 16, 1.0 + c) for c in range(280)]] + [[(r, 2.0)] for r in range(1, 20)] + [[(5, 3.0)]] * 13
    cases["dup_wrap"] = rows
    # empty rows and an empty tile row, rowA and colA not multiples of 16
    rows = []
    for i in range(53):
        if 16 <= i < 32 or i % 7 == 3:
            rows.append([])
        else:
            cols = sorted(set(rng.integers(0, 41, size=rng.integers(1, 6)).tolist()), reverse=bool(i % 2))
            rows.append([(c, float(rng.standard_normal())) for c in cols])
    cases["ragged"] = rows
    # one dense 16x16 tile (256 nonzeros) in file order that is not sorted
    rows = [[(c, float(r * 16 + c) + 0.5) for c in reversed(range(16))] for r in range(16)]
    cases["dense_tile"] = rows
    # wide: a few long rows spanning many tile columns
    rows = []
    for i in range(64):
        k = 1 + (200 if i % 17 == 0 else int(rng.integers(0, 4)))
        cols = rng.choice(600, size=k, replace=False)
        rows.append([(int(c), float(rng.standard_normal())) for c in cols])
    cases["wide"] = rows
    out = {}
    for name, rows in cases.items():
        ncol = 1 + max((c for r in rows for c, _ in r), default=0)
        if name == "ragged":
            ncol = 41
        rp = np.zeros(len(rows) + 1, np.int32)
        ci, va = [], []
        for i, r in enumerate(rows):
            for c, v in r:
                ci.append(c)
                va.append(v)
            rp[i + 1] = len(ci)
        out[name] = dict(m=len(rows), n=ncol, rowptr=rp, col=np.array(ci, np.int32),
                         val=np.array(va, np.float64))
    return out


def ref_case(R16, R8, csr):
    d = {"m": csr["m"], "n": csr["n"], "rowptr": csr["rowptr"], "col": csr["col"], "val": csr["val"]}
    for B, R in ((16, R16), (8, R8)):
        for trunc in (False, True):
            rowA = (csr["m"] // B) * B if trunc else csr["m"]
            if rowA == 0:
                continue
            key = f"b{B}_{'trunc' if trunc else 'full'}"
            t = R.tiles(rowA, csr["n"], csr["rowptr"], csr["col"], csr["val"])
            d[f"{key}_rowA"] = rowA
            for k in REF_ARRAYS:
                d[f"{key}_{k}"] = t[k]
            for k, v in t["offsets"].items():
                d[f"{key}_off_{k}"] = np.asarray(v)
            d[f"{key}_encode"] = R.encode(t["Tile_csr_Col"], 0)
    return d


def main():
    R16, R8 = Ref(16), Ref(8)
    for name, path in SHIPPED.items():
        rc, csr = R16.mmio(path)
        assert rc == 0, (name, rc)
        d = ref_case(R16, R8, csr)
        d["nnz"] = csr["nnz"]
        d["sym"] = csr["sym"]
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **d)
        if name in COPY_MTX:
            shutil.copyfile(path, os.path.join(HERE, f"{name}.mtx"))
        print("wrote", name, csr["m"], csr["nnz"])
    for name, csr in synthetic_cases().items():
        d = ref_case(R16, R8, csr)
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **d)
        print("wrote", name, csr["m"], len(csr["col"]))


if __name__ == "__main__":
    main()
