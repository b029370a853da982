"""Write golden vectors from the reference (ogauthe/frostspin 1.0.0) for what
only the reference defines:

* SU(2) block contents for representations with external degeneracies and
  mixed signatures (in-block row order, fusion-path order, spin-reversal signs;
  lie_group_symmetric_tensor.py:856-1109) and the blocks of three permutes of
  each tensor (lie_group_symmetric_tensor.py:1116-1134);
* O(2) blocks with 0odd / 0even / doublet irreps (o2_symmetric_tensor.py:118-352)
  and permutes of them;
* SimpleUpdate weights, logZ and bond representations after ``evolve``
  (simple_update.py:599-698, abstract_simple_update.py:94-147), U(1) and SU(2),
  with the reference's own checkpoint file of the final state;
* CTMRG trajectories: chi values, corner spectra and first-neighbour rdms over
  three iterations of the reference's AB/BA random cell at chi = 20
  (tests/test_ctmrg.py of the reference) and of Heisenberg environments built
  from the reference's simple-update checkpoints (ctm_renormalize.py:6-136),
  with the reference's CTMRG checkpoint files;
* the classical Ising example (examples/ising_classical_ctmrg.py) run at
  beta = 0.44, chi = 32, 200 iterations: energies, magnetization, the final Z2
  corner weights and edge tensor.

Run in the build container only (it imports the reference checkout, which
needs numba; tests/golden/numba_standin provides the four numba symbols the
reference uses as plain Python):

    python tests/golden/make_reference_fixtures.py

Outputs (data only, no reference source) land in tests/golden/ref_*.npz.
"""

import ast
import contextlib
import io
import os
import sys
import time

sys.dont_write_bytecode = True  # the reference checkout is read-only
HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = "/root/reference"
sys.path[:0] = [os.path.join(HERE, "numba_standin"), REFERENCE]

import numpy as np  # noqa: E402

from frostspin import (  # noqa: E402
    AsymmetricTensor,
    O2SymmetricTensor,
    SU2SymmetricTensor,
    U1SymmetricTensor,
)
from frostspin.ctmrg import SequentialCTMRG  # noqa: E402
from frostspin.simple_update import SimpleUpdate  # noqa: E402

SDS = np.array(
    [[0.25, 0.0, 0.0, 0.0], [0.0, -0.25, 0.5, 0.0], [0.0, 0.5, -0.25, 0.0], [0.0, 0.0, 0.0, 0.25]]
).reshape(2, 2, 2, 2)


def _out(name):
    return os.path.join(HERE, name)


def _put_tensor(data, prefix, t):
    """Reps, signature and blocks of a reference tensor (its own npz schema)."""
    data.update(t.get_data_dic(prefix=prefix))


# --------------------------------------------------------------------------------------
# (i) SU(2) and (ii) O(2) block layouts
# --------------------------------------------------------------------------------------
SU2_CASES = [
    # rank 3, one leg with external degeneracies
    {
        "row": [[[1, 2, 1], [1, 2, 3]], [[2, 1], [1, 2]]],
        "col": [[[1, 2, 1], [1, 2, 3]]],
        "sig": [False, True, True],
        "perms": [((1,), (0, 2)), ((2, 0), (1,)), ((0, 2), (1,))],
    },
    # rank 4, a chi-like leg (several multiplets of each spin), mixed signatures
    {
        "row": [[[3, 2, 2, 1], [1, 2, 3, 4]], [[1, 1], [1, 3]]],
        "col": [[[1, 2, 1], [1, 2, 3]], [[2, 1], [1, 2]]],
        "sig": [False, True, False, True],
        "perms": [((1, 3), (0, 2)), ((2,), (0, 3, 1)), ((3, 2, 0), (1,))],
    },
    # rank 6 site tensor (d, a | D, D, D, D) with degenerate virtual legs
    {
        "row": [[[1], [2]], [[1], [2]]],
        "col": [[[1, 1], [1, 2]], [[2, 1], [1, 3]], [[1, 1], [1, 2]], [[1, 1], [2, 3]]],
        "sig": [False, True, True, False, True, True],
        "perms": [((0, 2), (1, 3, 4, 5)), ((3, 4, 0), (5, 1, 2)), ((5,), (4, 3, 2, 1, 0))],
    },
]

O2_CASES = [
    {
        "row": [[[1, 1, 1, 1], [-1, 0, 1, 2]], [[2, 1, 1, 1], [-1, 0, 1, 2]], [[1, 2, 1, 1], [-1, 0, 1, 2]]],
        "col": [[[1, 1, 1, 1], [-1, 0, 2, 3]], [[2, 1, 1, 1], [-1, 0, 2, 3]]],
        "sig": [False, False, True, True, False],
        "perms": [((2, 1, 4), (0, 3)), ((0,), (1, 2, 3, 4)), ((4, 3), (2, 1, 0))],
    },
    {
        "row": [[[2, 1, 2], [-1, 0, 1]], [[1, 1, 1], [0, 1, 2]]],
        "col": [[[1, 2, 1], [-1, 0, 2]], [[1, 3], [0, 1]]],
        "sig": [True, False, True, False],
        "perms": [((1, 3), (0, 2)), ((3, 2, 0), (1,)), ((2,), (3, 1, 0))],
    },
]


def make_layouts(cls, cases, fname, seed):
    rng = np.random.default_rng(seed)
    data = {"n_cases": len(cases)}
    for ic, case in enumerate(cases):
        row = tuple(np.array(r, dtype=np.int64) for r in case["row"])
        col = tuple(np.array(r, dtype=np.int64) for r in case["col"])
        sig = np.array(case["sig"])
        t = cls.random(row, col, signature=sig, rng=rng)
        t /= t.norm()
        dense = t.toarray()
        back = cls.from_array(dense, row, col, signature=sig)
        assert all(np.abs(b1 - b2).max() < 1e-13 for b1, b2 in zip(back.blocks, t.blocks, strict=True))
        p = f"case{ic}"
        data[p + "_dense"] = dense
        data[p + "_n_perms"] = len(case["perms"])
        _put_tensor(data, p + "_t", t)
        for ip, (ra, ca) in enumerate(case["perms"]):
            tp = t.permute(ra, ca)
            data[f"{p}_perm{ip}_row_axes"] = np.array(ra, dtype=np.int64)
            data[f"{p}_perm{ip}_col_axes"] = np.array(ca, dtype=np.int64)
            _put_tensor(data, f"{p}_perm{ip}", tp)
            tt = tp.transpose()
            _put_tensor(data, f"{p}_perm{ip}_T", tt)
    np.savez_compressed(_out(fname), **data)
    print(fname, "written")


# --------------------------------------------------------------------------------------
# (iii) SimpleUpdate, (iv) CTMRG
# --------------------------------------------------------------------------------------
def _heisenberg(sym):
    if sym == "U1":
        reps = (np.array([1, -1], dtype=np.int8),) * 2
        return U1SymmetricTensor.from_array(SDS, reps, reps)
    reps = (np.array([[1], [2]]),) * 2
    return SU2SymmetricTensor.from_array(SDS, reps, reps)


def _smallest_weight(su):
    """Smallest kept weight over the bonds (a cut that keeps tiny values would
    be fragile)."""
    return min(float(np.abs(w).min()) for w in su.get_weights())


SU_CASES = [
    # symmetry, D, tau, beta
    ("U1", 4, 0.01, 0.2),
    ("SU2", 4, 0.01, 0.2),
    ("U1", 6, 0.01, 0.3),
    ("SU2", 7, 0.01, 0.3),
    ("U1", 4, 0.01, 1.0),
    ("SU2", 4, 0.01, 1.0),
]


def make_simple_update():
    data = {"n_cases": len(SU_CASES)}
    for ic, (sym, D, tau, beta) in enumerate(SU_CASES):
        su = SimpleUpdate.square_lattice_first_neighbor(_heisenberg(sym), D, tau, degen_ratio=0.99, rcutoff=1e-10)
        su.evolve(beta)
        p = f"case{ic}"
        data[p + "_symmetry"] = sym
        data[p + "_D"] = D
        data[p + "_tau"] = tau
        data[p + "_beta"] = beta
        data[p + "_beta_reached"] = su.beta  # whole steps of 4 tau (abstract_simple_update.py:110)
        data[p + "_logZ"] = su.logZ
        data[p + "_Dmax"] = su.Dmax
        for ib, (w, r) in enumerate(zip(su.get_weights(), su.get_bond_representations(), strict=True)):
            data[f"{p}_weights_{ib}"] = np.asarray(w)
            data[f"{p}_bond_rep_{ib}"] = r
        su.save_to_file(_out(f"ref_su_checkpoint_{ic}.npz"))
        print("simple update", sym, D, beta, "logZ", su.logZ, "min weight", _smallest_weight(su))
    np.savez_compressed(_out("ref_simple_update.npz"), **data)


def _rdms(ctm):
    r21, r12 = ctm.compute_rdm_1st_neighbor_cell()
    return [r.toarray(as_matrix=True) for r in r21], [r.toarray(as_matrix=True) for r in r12]


def _corner_spectra(ctm):
    out = []
    for x, y in ctm.site_coords:
        for getter in (ctm.get_C1, ctm.get_C2, ctm.get_C3, ctm.get_C4):
            m = getter(x, y).toarray(as_matrix=True)
            out.append(np.linalg.svd(m, compute_uv=False))
    return out


def _run_ctm(data, p, ctm, n_iter, save_name):
    chis = []
    for it in range(n_iter):
        ctm.iterate()
        chis.append(list(map(int, ctm.chi_values)))
        r21, r12 = _rdms(ctm)
        for i, (a, b) in enumerate(zip(r21, r12, strict=True)):
            data[f"{p}_it{it}_rdm2x1_{i}"] = a
            data[f"{p}_it{it}_rdm1x2_{i}"] = b
        for k, s in enumerate(_corner_spectra(ctm)):
            data[f"{p}_it{it}_corner_sv_{k}"] = s
    data[p + "_chi_values"] = np.array(
        [c + [0] * (max(map(len, chis)) - len(c)) for c in chis], dtype=np.int64
    )
    data[p + "_n_iter"] = n_iter
    if save_name is not None:
        ctm.save_to_file(_out(save_name))


def _random_cell(seed):
    """Tensors of the reference's tests/test_ctmrg.py (seed 42 there)."""
    rng = np.random.default_rng(seed)
    rp = np.array([1, -1], dtype=np.int8)
    ra = np.array([1, -1, 2], dtype=np.int8)
    ru = np.array([1, 1, -1, 0, 2, -2], dtype=np.int8)
    rr = np.array([2, 0, 1, -1, 0, 1, -1], dtype=np.int8)
    rd = np.array([1, 1, -1, -2], dtype=np.int8)
    rl = np.array([1, -2, 1, 0, 0], dtype=np.int8)
    sA = np.array([False, False, True, True, True, True])
    sB = np.array([False, False, False, False, False, False])
    reps_A = (rp, ra, ru, rr, rd, rl)
    reps_B = (rp, ra, rd, rl, ru, rr)
    A = U1SymmetricTensor.random(reps_A[:2], reps_A[2:], rng=rng, signature=sA)
    B = U1SymmetricTensor.random(reps_B[:2], reps_B[2:], rng=rng, signature=sB)
    dA = tuple(np.array([t.size]) for t in reps_A)
    dB = tuple(np.array([t.size]) for t in reps_B)
    A_as = AsymmetricTensor.from_array(A.toarray(), dA[:2], dA[2:], signature=sA)
    B_as = AsymmetricTensor.from_array(B.toarray(), dB[:2], dB[2:], signature=sB)
    return (A, B), (A_as, B_as)


CTM_PARAMS = {"degen_ratio": 0.99, "cutoff": 1e-10}
# (simple-update case, name): D = 6 / 7 at beta = 0.3, D = 4 at beta = 1
HEIS_CTM_CASES = ((2, "heis_u1_d6"), (3, "heis_su2_d7"), (4, "heis_u1_b1"), (5, "heis_su2_b1"))


def make_ctmrg():
    data = {}
    t0 = time.time()
    # (a) random U(1) cell of the reference's tests, default parameters
    u1, _ = _random_cell(42)
    ctm = SequentialCTMRG.from_elementary_tensors("AB\nBA", u1, 20)
    _run_ctm(data, "random_u1", ctm, 3, "ref_ctm_random_u1.npz")
    print("random cell", data["random_u1_chi_values"].tolist(), f"{time.time() - t0:.1f}s")
    # (b) Heisenberg environments from the reference's simple-update checkpoints
    for ic, name in HEIS_CTM_CASES:
        t0 = time.time()
        su = SimpleUpdate.load_from_file(_out(f"ref_su_checkpoint_{ic}.npz"))
        chi = 24
        ctm = SequentialCTMRG.from_elementary_tensors("AB\nBA", su.get_tensors(), chi, **CTM_PARAMS)
        _run_ctm(data, name, ctm, 3, f"ref_ctm_{name}.npz")
        data[name + "_su_case"] = ic
        data[name + "_chi"] = chi
        print(name, data[name + "_chi_values"].tolist(), f"{time.time() - t0:.1f}s")
    np.savez_compressed(_out("ref_ctmrg.npz"), **data)


# --------------------------------------------------------------------------------------
# (v) classical Ising example at the plain-method configuration
# --------------------------------------------------------------------------------------
ISING_PARAMS = {"beta": 0.44, "chi_target": 32, "n_iter": 200}


def make_ising():
    """Run examples/ising_classical_ctmrg.py unchanged except for its three
    top-level parameters, and keep its results."""
    path = os.path.join(REFERENCE, "examples", "ising_classical_ctmrg.py")
    with open(path, encoding="utf8") as fin:
        tree = ast.parse(fin.read(), filename=path)
    seen = set()
    for node in tree.body:
        if isinstance(node, ast.Assign) and len(node.targets) == 1 and isinstance(node.targets[0], ast.Name):
            name = node.targets[0].id
            if name in ISING_PARAMS and name not in seen:
                node.value = ast.Constant(ISING_PARAMS[name])
                seen.add(name)
    assert seen == set(ISING_PARAMS), seen
    code = compile(ast.fix_missing_locations(tree), path, "exec")
    glb = {"__name__": "ising_example", "__file__": path}
    t0 = time.time()
    with contextlib.redirect_stdout(io.StringIO()):
        exec(code, glb)  # noqa: S102
    s, T = glb["s"], glb["T"]  # Z2 run is the last one
    data = {
        "beta": ISING_PARAMS["beta"],
        "chi": ISING_PARAMS["chi_target"],
        "n_iter": ISING_PARAMS["n_iter"],
        "seed": 42,
        "energy_trivial": float(glb["energy"]),
        "magnetization_trivial": float(glb["m"]),
        "energy_Z2": float(glb["energy_Z2"]),
        "magnetization_Z2": float(glb["m_Z2"]),
        "s_block_irreps": np.asarray(s.block_irreps),
        "T_dense": T.toarray(),
    }
    for i, b in enumerate(s.diagonal_blocks):
        data[f"s_block_{i}"] = np.asarray(b)
    _put_tensor(data, "T", T)
    np.savez_compressed(_out("ref_ising.npz"), **data)
    print("ising", data["energy_trivial"], data["energy_Z2"], f"{time.time() - t0:.1f}s")


def main():
    which = set(sys.argv[1:]) or {"layouts", "su", "ctm", "ising"}
    if "layouts" in which:
        make_layouts(SU2SymmetricTensor, SU2_CASES, "ref_su2_layouts.npz", 7)
        make_layouts(O2SymmetricTensor, O2_CASES, "ref_o2_layouts.npz", 11)
    if "su" in which:
        make_simple_update()
    if "ctm" in which:
        make_ctmrg()
    if "ising" in which:
        make_ising()


if __name__ == "__main__":
    main()
