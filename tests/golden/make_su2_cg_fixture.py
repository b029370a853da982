"""Write tests/golden/su2_cg_reference.npz: the reference's SU(2)
Clebsch-Gordan table (frostspin/groups/clebsch-gordan_data/
data_su2_clebsch-gordan.json, a data file of the reference) restricted to
irrep dimensions d1, d2 <= 6, stored sparse like the reference's npz format
(i_d1_d2_d3 flat indices, c_d1_d2_d3 coefficients).

Run in the build container (reads the reference checkout):
    python tests/golden/make_su2_cg_fixture.py
"""

import json
import os

import numpy as np

SRC = "/root/reference/frostspin/groups/clebsch-gordan_data/data_su2_clebsch-gordan.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "su2_cg_reference.npz")
MAX_DIM = 6


def main():
    with open(SRC, encoding="utf8") as fin:
        data = json.load(fin)
    out = {}
    for key, val in data.items():
        if key.startswith("!"):
            continue
        d1, d2, d3 = (int(x) for x in key.split(","))
        if d1 > MAX_DIM or d2 > MAX_DIM:
            continue
        inds, coeff = val
        out[f"i_{d1}_{d2}_{d3}"] = np.asarray(inds, dtype=np.int64)
        out[f"c_{d1}_{d2}_{d3}"] = np.asarray(coeff, dtype=np.float64)
    np.savez_compressed(OUT, **out)
    print(OUT, len(out) // 2, "tensors")


if __name__ == "__main__":
    main()
