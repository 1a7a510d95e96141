"""Writes tests/golden/small4.tns: a 4-mode coordinate tensor in SPLATT's
text format (one nonzero per line, 1-based indices then the value), from a
seeded numpy generator.  Holes in every mode (empty slices), 40 repeated
coordinates (kept as separate nonzeros by splatt_csf_load), a comment line
and a blank line.  Values are written with 17 significant digits, so they
parse back exactly."""
from pathlib import Path

import numpy as np


def main() -> None:
    rng = np.random.default_rng(20241204)
    dims = [30, 25, 20, 12]
    nnz = 3000
    inds = np.stack([rng.integers(0, d, nnz) for d in dims])
    for m, d in enumerate(dims):  # indices = 3 mod 7 never occur
        inds[m] = np.where(inds[m] % 7 == 3, (inds[m] + 1) % d, inds[m])
    inds[:, 1000:1040] = inds[:, :40]
    vals = rng.standard_normal(nnz)
    out = Path(__file__).with_name("small4.tns")
    with out.open("w") as f:
        f.write("# 4-mode test tensor (tests/golden/make_tns.py)\n\n")
        for j in range(nnz):
            f.write(" ".join(str(int(inds[m, j]) + 1) for m in range(4)))
            f.write(f" {vals[j]:.17g}\n")


if __name__ == "__main__":
    main()
