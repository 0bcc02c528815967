"""Write golden vectors from the reference's pure-Python operator and layout code.

Run in the build container only (it reads /root/reference, which does not
exist on GPU machines):

    python tests/golden/make_golden.py

Only modules that import nothing beyond numpy are loaded:
slowquant/unitary_coupled_cluster/{fermionic_operator,operators,util}.py.
They pin what no mathematical identity fixes: the reference's exact
normal-ordered strings (incl. its folding of inactive orbitals) and its ansatz
gate lists.  Outputs are data only (JSON), committed beside this script.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _strings(op):
    return [[[[int(i), bool(d)] for i, d in key], float(c)] for key, c in op.operators.items()]


def main() -> None:
    sys.path.insert(0, REF)
    from slowquant.unitary_coupled_cluster import operators as O  # noqa: E402
    from slowquant.unitary_coupled_cluster import util as U  # noqa: E402

    ops = {}
    ops["G1(0,2)"] = _strings(O.G1(0, 2, True))
    ops["G1(1,5)"] = _strings(O.G1(1, 5, True))
    ops["G2(0,1,2,3)"] = _strings(O.G2(0, 1, 2, 3, True))
    ops["G2(0,3,4,7)"] = _strings(O.G2(0, 3, 4, 7, True))
    ops["G2(2,1,6,5)"] = _strings(O.G2(2, 1, 6, 5, True))
    ops["G3(0,1,2,4,5,7)"] = _strings(O.G3(0, 1, 2, 4, 5, 7, True))
    ops["G1_sa(0,2)"] = _strings(O.G1_sa(0, 2, True))
    ops["G2_sa(0,0,2,2,1)"] = _strings(O.G2_sa(0, 0, 2, 2, 1, True))
    ops["G2_sa(0,0,2,3,2)"] = _strings(O.G2_sa(0, 0, 2, 3, 2, True))
    ops["G2_sa(0,1,2,2,3)"] = _strings(O.G2_sa(0, 1, 2, 2, 3, True))
    ops["G2_sa(0,1,2,3,4)"] = _strings(O.G2_sa(0, 1, 2, 3, 4, True))
    ops["G2_sa(0,1,2,3,5)"] = _strings(O.G2_sa(0, 1, 2, 3, 5, True))
    ops["G2_sa(1,0,3,2,4)"] = _strings(O.G2_sa(1, 0, 3, 2, 4, True))
    ops["Epq(1,2)*Epq(2,0)"] = _strings(O.Epq(1, 2) * O.Epq(2, 0))
    ops["epqrs(0,1,1,2)"] = _strings(O.epqrs(0, 1, 1, 2))

    # Hamiltonian with one inactive and two active orbitals, folded.
    rng = np.random.default_rng(11)
    n = 3
    A = rng.normal(size=(n, n))
    h = A + A.T
    B = rng.normal(size=(2, n, n))
    B = B + B.transpose(0, 2, 1)
    g = np.einsum("Lpq,Lrs->pqrs", B, B)
    H = O.hamiltonian_0i_0a(h, g, 1, 2)
    ham = {
        "h": h.tolist(),
        "g": g.reshape(-1).tolist(),
        "n_inact": 1,
        "n_act": 2,
        "strings": _strings(H),
        "folded": _strings(H.get_folded_operator(1, 2, 0)),
    }

    layouts = {}

    def dump(name, s):
        layouts[name] = {
            "types": list(s.excitation_operator_type),
            "indices": [list(map(int, t)) for t in s.excitation_indices],
            "n_params": int(s.n_params),
            "grad_param_R": dict(getattr(s, "grad_param_R", {})),
            "param_names": list(getattr(s, "param_names", [])),
        }

    for n_orb, opts in (
        (4, {"n_layers": 1, "do_tups": True}),
        (5, {"n_layers": 2, "do_tups": True}),
        (6, {"n_layers": 2, "do_tups": True, "skip_last_singles": True}),
        (2, {"n_layers": 1, "do_tups": True, "skip_last_singles": True}),
        (6, {"n_layers": 2, "do_qnp": True}),
        (12, {"n_layers": 4, "do_tups": True}),
    ):
        s = U.UpsStructure()
        s.create_tiled(n_orb, dict(opts))
        dump(f"tiled_{n_orb}_{sorted(opts.items())}", s)

    occ, unocc = [0, 1], [2, 3]
    occ_s, unocc_s = [0, 1, 2, 3], [4, 5, 6, 7]
    for opts in (
        {"n_layers": 1, "S": True, "D": True},
        {"n_layers": 2, "SAS": True, "SAD": True},
        {"n_layers": 1, "SAGS": True, "GpD": True},
        {"n_layers": 1, "GS": True, "GD": True, "pD": True},
        {"n_layers": 1, "T": True, "Q": True},
    ):
        s = U.UpsStructure()
        s.create_fUCC(occ, unocc, occ_s, unocc_s, 4, dict(opts))
        dump(f"fucc_{sorted(opts.items())}", s)
    s = U.UpsStructure()
    s.create_fUCC([0, 1, 2, 3], [4, 5, 6, 7], list(range(8)), list(range(8, 16)), 8, {"n_layers": 1, "SAS": True, "SAD": True})
    dump("fucc_8_8_SAS_SAD", s)
    for opts in ({"n_layers": 1, "D": True}, {"n_layers": 1, "GpD": True}, {"n_layers": 1, "pD": True}):
        s = U.UpsStructure()
        s.create_SDSfUCC(occ, unocc, occ_s, unocc_s, 4, dict(opts))
        dump(f"sds_{sorted(opts.items())}", s)
    u = U.UccStructure()
    u.add_sa_singles(occ, unocc)
    u.add_sa_doubles(occ, unocc)
    u.add_triples(occ_s, unocc_s)
    u.add_quadruples(occ_s, unocc_s)
    dump("ucc_SDTQ_4_4", u)
    u = U.UccStructure()
    u.add_sa_singles([0, 1, 2], [3, 4, 5])
    u.add_sa_doubles([0, 1, 2], [3, 4, 5])
    u.add_triples([0, 1, 2, 3, 4, 5], [6, 7, 8, 9, 10, 11])
    dump("ucc_SDT_6_6", u)

    with open(os.path.join(HERE, "operators.json"), "w") as f:
        json.dump({"operators": ops, "hamiltonian": ham}, f)
    with open(os.path.join(HERE, "layouts.json"), "w") as f:
        json.dump(layouts, f)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
