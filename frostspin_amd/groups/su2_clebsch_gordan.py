"""SU(2) Clebsch-Gordan tensors.

An irrep is labelled by its dimension d = 2j + 1. The tensor for (d1, d2, d3)
has shape (d1, d2, d3); index i of an irrep of spin j stands for m = j - i
(descending weights), Condon-Shortley phases. This is the convention of
frostspin/groups/su2_clebsch_gordan.py:8-32, which stores the same tensors
(computed with sympy) in a JSON/npz table limited to d <= 10.

Here the coefficients are computed on demand from Racah's closed form in exact
rational arithmetic, so any dimension is available without a table. Tables in
the reference's JSON or npz format can still be loaded (load_su2_cg) and take
precedence over the computed values.
"""

import json
from fractions import Fraction
from functools import lru_cache
from math import factorial, isqrt

import numpy as np


def _sqrt_fraction(x):
    """float(sqrt(x)) for a non-negative Fraction, accurate to ~1 ulp."""
    if x == 0:
        return 0.0
    p, q = x.numerator, x.denominator
    bits = 128
    r = isqrt((p << (2 * bits)) * q)  # sqrt(p q) * 2^bits
    return float(Fraction(r, q << bits))


def _cg_coefficient(d1, i1, d2, i2, d3, i3):
    """<j1 m1; j2 m2 | j3 m3> with j = (d-1)/2 and m = j - i (Racah formula)."""
    # doubled quantum numbers keep everything integer
    J1, J2, J3 = d1 - 1, d2 - 1, d3 - 1
    M1, M2, M3 = J1 - 2 * i1, J2 - 2 * i2, J3 - 2 * i3
    if M1 + M2 != M3:
        return 0.0
    if (J1 + J2 + J3) % 2 or J3 < abs(J1 - J2) or J3 > J1 + J2:
        return 0.0
    a = (J1 + J2 - J3) // 2
    b = (J1 - J2 + J3) // 2
    c = (-J1 + J2 + J3) // 2
    s = (J1 + J2 + J3) // 2 + 1
    j1p, j1m = (J1 + M1) // 2, (J1 - M1) // 2
    j2p, j2m = (J2 + M2) // 2, (J2 - M2) // 2
    j3p, j3m = (J3 + M3) // 2, (J3 - M3) // 2
    pref = Fraction((J3 + 1) * factorial(a) * factorial(b) * factorial(c), factorial(s))
    pref *= (
        factorial(j1p) * factorial(j1m) * factorial(j2p) * factorial(j2m) * factorial(j3p) * factorial(j3m)
    )
    tot = Fraction(0)
    # k runs over all values keeping every factorial argument >= 0
    kmin = max(0, (J2 - J3 - M1) // 2, (J1 - J3 + M2) // 2)
    kmax = min(a, j1m, j2p)
    for k in range(kmin, kmax + 1):
        den = (
            factorial(k)
            * factorial(a - k)
            * factorial(j1m - k)
            * factorial(j2p - k)
            * factorial((J3 - J2 + M1) // 2 + k)
            * factorial((J3 - J1 - M2) // 2 + k)
        )
        tot += Fraction((-1) ** k, den)
    if tot == 0:
        return 0.0
    mag = _sqrt_fraction(pref * tot * tot)
    return mag if tot > 0 else -mag


@lru_cache(maxsize=None)
def _computed_cg(d1, d2, d3):
    p = np.zeros((d1, d2, d3))
    for i1 in range(d1):
        for i2 in range(d2):
            # m3 = m1 + m2 fixes i3
            twice_i3 = (d3 - 1) - ((d1 - 1 - 2 * i1) + (d2 - 1 - 2 * i2))
            if twice_i3 % 2 or not 0 <= twice_i3 // 2 < d3:
                continue
            i3 = twice_i3 // 2
            p[i1, i2, i3] = _cg_coefficient(d1, i1, d2, i2, d3, i3)
    p.flags.writeable = False
    return p


_loaded_table = {}


def clebsch_gordan(d1, d2, d3):
    """Read-only CG tensor (d1, d2, d3); d3 must be in |d1-d2|+1 .. d1+d2-1 step 2."""
    d1, d2, d3 = int(d1), int(d2), int(d3)
    if d3 < abs(d1 - d2) + 1 or d3 > d1 + d2 - 1 or (d1 + d2 + d3) % 2 == 0:
        raise ValueError(f"irrep {d3} does not appear in {d1} x {d2}")
    t = _loaded_table.get((d1, d2, d3))
    return t if t is not None else _computed_cg(d1, d2, d3)


def spin_reversal(d):
    """Matrix Z of the spin-reversal map applied on a leg whose signature is
    flipped: (Z v)[i] = s_i v[d-1-i] with s = (-1, 1, -1, ...) for even d and
    (1, -1, 1, ...) for odd d (su2_symmetric_tensor.py:76-91)."""
    d = int(d)
    signs = (np.arange(d % 2, d + d % 2) % 2) * 2 - 1
    z = np.zeros((d, d))
    z[np.arange(d), d - 1 - np.arange(d)] = signs
    return z


def _read_table(path):
    path = str(path)
    table = {}
    if path.endswith(".json"):
        with open(path, encoding="utf8") as fin:
            data = json.load(fin)
        for key, val in data.items():
            if key.startswith("!"):
                continue
            sh = tuple(int(x) for x in key.split(","))
            inds, coeff = val
            flat = np.zeros((sh[0] * sh[1] * sh[2],))
            flat[np.asarray(inds, dtype=np.int64)] = coeff
            table[sh] = flat.reshape(sh)
    elif path.endswith(".npz"):
        with np.load(path) as data:  # allow_pickle stays False
            for key in data.files:
                if key[0] != "i":
                    continue
                sh = tuple(int(x) for x in key[2:].split("_"))
                flat = np.zeros((sh[0] * sh[1] * sh[2],))
                flat[data[key]] = data["c" + key[1:]]
                table[sh] = flat.reshape(sh)
    else:
        msg = f"Invalid file extension: must be json or npz, got {path.rsplit('.', 1)[-1]}"
        raise RuntimeError(msg)
    return table


def load_su2_cg(cg_file=None, *, verbosity=0):
    """Load a Clebsch-Gordan table in the format of su2_clebsch_gordan.py:35-109
    (JSON sparse lists or npz 'i_/c_' arrays). Loaded tensors replace the computed
    ones; dimensions outside the table are still computed. cg_file None takes
    config["SU2_CG_file"] (the --frostspin-SU2-CG-file option, as in the
    reference); when that is unset too, the computed coefficients are used
    alone. A file that cannot be read raises FileNotFoundError. Returns the
    loaded table."""
    from ..config import config

    _loaded_table.clear()
    if cg_file is None:
        cg_file = config["SU2_CG_file"]
    if cg_file is None:
        return {}
    cg_file = str(cg_file)
    try:
        table = _read_table(cg_file)
    except FileNotFoundError:
        msg = f"Error: cannot read SU(2) Clebsch-Gordan coefficient savefile {cg_file}"
        raise FileNotFoundError(msg) from None
    for k, v in table.items():
        v.flags.writeable = False
        _loaded_table[k] = v
    if verbosity > 0:
        print("Loaded SU(2) Clebsch-Gordan coefficients from", cg_file)
    return dict(table)


def save_su2_cg(savefile, max_spin_dim):
    """Write computed tensors for all dimensions <= max_spin_dim in the reference's
    JSON or npz format (su2_clebsch_gordan.py:35-160)."""
    savefile = str(savefile)
    tensors = {}
    for d1 in range(1, max_spin_dim + 1):
        for d2 in range(1, max_spin_dim + 1):
            for d3 in range(abs(d1 - d2) + 1, d1 + d2, 2):
                tensors[d1, d2, d3] = _computed_cg(d1, d2, d3)
    if savefile.endswith(".npz"):
        data = {}
        for (d1, d2, d3), p in tensors.items():
            inds = np.flatnonzero(p)
            data[f"i_{d1}_{d2}_{d3}"] = inds
            data[f"c_{d1}_{d2}_{d3}"] = p.flat[inds]
        np.savez_compressed(savefile, **data)
    elif savefile.endswith(".json"):
        data = {
            "!!info": ["SU(2) Clebsch-Gordan tensors, irreps labelled by dimension, flat C order"],
            "!maximal_spin_dimension": int(max_spin_dim),
        }
        for (d1, d2, d3), p in tensors.items():
            inds = np.flatnonzero(p)
            data[f"{d1},{d2},{d3}"] = [inds.tolist(), p.flat[inds].tolist()]
        with open(savefile, "w", encoding="utf8") as out:
            json.dump(data, out, indent=2, sort_keys=True)
    else:
        raise RuntimeError("Invalid file extension: must be json or npz")


# a table named on the command line (--frostspin-SU2-CG-file) is read once at
# import, like the reference's class attribute (su2_symmetric_tensor.py:140)
load_su2_cg()
