"""CPU tests of the host-side mirror of the reference interface and of the
C-ABI library (load + exported symbols; no compute calls without a GPU)."""

from __future__ import annotations

import ctypes
import json
import os
import re

import numpy as np
import pytest

from slowquant_amd import fermionic_operator as FO
from slowquant_amd import operators as OP
from slowquant_amd.util import UccStructure, UpsStructure

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _keyed(strings):
    return {tuple((int(i), bool(d)) for i, d in k): c for k, c in strings}


def _golden_layouts():
    return json.load(open(os.path.join(GOLD, "layouts.json")))


def _cmp(d, s):
    assert d["types"] == list(s.excitation_operator_type)
    assert d["indices"] == [list(t) for t in s.excitation_indices]
    assert d["n_params"] == s.n_params
    assert d["grad_param_R"] == dict(getattr(s, "grad_param_R", {}))
    assert d["param_names"] == list(getattr(s, "param_names", []))


TILED = [
    (4, {"n_layers": 1, "do_tups": True}),
    (5, {"n_layers": 2, "do_tups": True}),
    (6, {"n_layers": 2, "do_tups": True, "skip_last_singles": True}),
    (2, {"n_layers": 1, "do_tups": True, "skip_last_singles": True}),
    (6, {"n_layers": 2, "do_qnp": True}),
    (12, {"n_layers": 4, "do_tups": True}),
]


@pytest.mark.parametrize("n,opts", TILED)
def test_tiled_layouts_golden(n, opts):
    s = UpsStructure()
    s.create_tiled(n, dict(opts))
    _cmp(_golden_layouts()[f"tiled_{n}_{sorted(opts.items())}"], s)


FUCC = [
    {"n_layers": 1, "S": True, "D": True},
    {"n_layers": 2, "SAS": True, "SAD": True},
    {"n_layers": 1, "SAGS": True, "GpD": True},
    {"n_layers": 1, "GS": True, "GD": True, "pD": True},
    {"n_layers": 1, "T": True, "Q": True},
]


@pytest.mark.parametrize("opts", FUCC)
def test_fucc_layouts_golden(opts):
    s = UpsStructure()
    s.create_fUCC([0, 1], [2, 3], [0, 1, 2, 3], [4, 5, 6, 7], 4, dict(opts))
    _cmp(_golden_layouts()[f"fucc_{sorted(opts.items())}"], s)


def test_other_layouts_golden():
    L = _golden_layouts()
    s = UpsStructure()
    s.create_fUCC([0, 1, 2, 3], [4, 5, 6, 7], list(range(8)), list(range(8, 16)), 8, {"n_layers": 1, "SAS": True, "SAD": True})
    _cmp(L["fucc_8_8_SAS_SAD"], s)
    assert s.n_params == 152  # SAfUCCSD at (8e,8o): 16 singles + 136 doubles
    for opts in ({"n_layers": 1, "D": True}, {"n_layers": 1, "GpD": True}, {"n_layers": 1, "pD": True}):
        s = UpsStructure()
        s.create_SDSfUCC([0, 1], [2, 3], [0, 1, 2, 3], [4, 5, 6, 7], 4, dict(opts))
        _cmp(L[f"sds_{sorted(opts.items())}"], s)
    u = UccStructure()
    u.add_sa_singles([0, 1], [2, 3])
    u.add_sa_doubles([0, 1], [2, 3])
    u.add_triples([0, 1, 2, 3], [4, 5, 6, 7])
    u.add_quadruples([0, 1, 2, 3], [4, 5, 6, 7])
    _cmp(L["ucc_SDTQ_4_4"], u)


def test_layout_option_errors():
    with pytest.raises(ValueError):
        UpsStructure().create_tiled(4, {"n_layers": 1})
    with pytest.raises(ValueError):
        UpsStructure().create_tiled(4, {"n_layers": 1, "do_tups": True, "bogus": 1})
    with pytest.raises(ValueError):
        UpsStructure().create_fUCC([0], [1], [0, 1], [2, 3], 2, {"n_layers": 1})


def test_product_operator_algebra_matches_golden():
    data = json.load(open(os.path.join(GOLD, "operators.json")))
    for name, strings in data["operators"].items():
        f, args = name.split("(", 1)
        if "*" in name:
            mine = OP.Epq(1, 2) * OP.Epq(2, 0)
        elif f == "epqrs":
            mine = OP.epqrs(0, 1, 1, 2)
        else:
            ints = [int(x) for x in args.rstrip(")").split(",")]
            mine = getattr(OP, f)(*ints, True)
        ref = _keyed(strings)
        assert set(ref) == set(mine.operators), name
        assert all(abs(ref[k] - mine.operators[k]) < 1e-15 for k in ref), name
    ham = data["hamiltonian"]
    h = np.array(ham["h"])
    g = np.array(ham["g"]).reshape(3, 3, 3, 3)
    H = OP.hamiltonian_0i_0a(h, g, 1, 2)
    assert max(abs(v - H.operators[k]) for k, v in _keyed(ham["strings"]).items()) < 1e-13
    Hf = H.get_folded_operator(1, 2, 0)
    ref_f = _keyed(ham["folded"])
    assert set(ref_f) == set(Hf.operators)
    assert max(abs(v - Hf.operators[k]) for k, v in ref_f.items()) < 1e-13


def test_folded_integrals_reproduce_reference_folding():
    """e_core + h_eff + g_act rebuild exactly the reference's folded operator."""
    ham = json.load(open(os.path.join(GOLD, "operators.json")))["hamiltonian"]
    h = np.array(ham["h"])
    g = np.array(ham["g"]).reshape(3, 3, 3, 3)
    e_core, h_eff, g_act = OP.folded_integrals(h, g, 1, 2)
    rebuilt = OP.hamiltonian_0i_0a(h_eff, g_act, 0, 2)
    if abs(e_core) > 1e-14:
        rebuilt = rebuilt + FO.FermionicOperator({(): e_core})
    ref = _keyed(ham["folded"])
    keys = set(ref) | set(rebuilt.operators)
    assert max(abs(ref.get(k, 0.0) - rebuilt.operators.get(k, 0.0)) for k in keys) < 1e-12


def test_ci_indexing_matches_oracle():
    from oracle.ci import get_indexing as oracle_indexing

    from slowquant_amd.ci_spaces import get_indexing

    for n, na, nb in [(4, 2, 2), (6, 3, 3), (5, 1, 3)]:
        ci = get_indexing(0, n, 0, na, nb)
        ref = oracle_indexing(0, n, 0, na, nb)
        assert np.array_equal(ci.idx2det, ref.idx2det)
        for i in range(0, len(ref.idx2det), 7):
            assert ci.det2idx[int(ref.idx2det[i])] == i
        assert len(ci.det2idx) == len(ref.idx2det)
        with pytest.raises(KeyError):
            ci.det2idx[0]


def test_density_matrix_functions_match_elementwise_rules():
    from slowquant_amd.density_matrix import RDM1, RDM2, get_electronic_energy, get_orbital_gradient

    rng = np.random.default_rng(0)
    ni, na, nv = 2, 2, 1
    N = ni + na + nv
    r1 = rng.normal(size=(na, na))
    r1 = r1 + r1.T
    r2 = rng.normal(size=(na,) * 4)
    h = rng.normal(size=(N, N))
    g = rng.normal(size=(N,) * 4)
    o = ni + na
    # scalar restatement of density_matrix.py:139-178 and :181-230
    E = sum(h[p, q] * RDM1(p, q, ni, na, r1) for p in range(o) for q in range(o))
    E += sum(0.5 * g[p, q, r, s] * RDM2(p, q, r, s, ni, na, r1, r2)
             for p in range(o) for q in range(o) for r in range(o) for s in range(o))
    assert abs(E - get_electronic_energy(h, g, ni, na, r1, r2)) < 1e-12
    kidx = [(p, q) for p in range(N) for q in range(p + 1, N)]
    ref = np.zeros(len(kidx))
    for i, (m, n) in enumerate(kidx):
        for p in range(o):
            ref[i] += 2 * h[n, p] * RDM1(m, p, ni, na, r1) - 2 * h[p, m] * RDM1(p, n, ni, na, r1)
            for q in range(o):
                for r in range(o):
                    ref[i] += (g[n, p, q, r] * RDM2(m, p, q, r, ni, na, r1, r2)
                               - g[p, m, q, r] * RDM2(p, n, q, r, ni, na, r1, r2)
                               - g[m, p, q, r] * RDM2(n, p, q, r, ni, na, r1, r2)
                               + g[p, n, q, r] * RDM2(p, m, q, r, ni, na, r1, r2))
    assert np.abs(ref - get_orbital_gradient(h, g, kidx, ni, na, r1, r2)).max() < 1e-11


def _header_functions():
    text = open(os.path.join(ROOT, "include", "sqhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    from slowquant_amd import _native

    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libsqhip.so not built")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [f for f in _header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    # every binding in _native refers to a declared symbol
    declared = set(_header_functions())
    assert set(_native.exported_symbols()) <= declared
    # pure host queries work without a GPU
    lib.sq_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.sq_version()


def test_wavefunction_construction_on_cpu():
    from slowquant_amd.synthetic import SyntheticIntegrals, synthetic_hamiltonian
    from slowquant_amd.ups_wavefunction import WaveFunctionUPS

    h, g = synthetic_hamiltonian(12, 24, np.random.default_rng(0))
    gen = SyntheticIntegrals(h, g, 12)
    wf = WaveFunctionUPS((12, 12), np.eye(12), gen, "tUPS", {"n_layers": 4})
    assert wf.ups_layout.n_params == 132
    assert wf.num_det == 853776
    assert wf.num_inactive_orbs == 0 and wf.num_virtual_orbs == 0
    assert wf.csf_coeffs[0] == 1.0
    assert len(wf.kappa) == 0
    with pytest.raises(ValueError):
        wf.thetas = [0.0] * 3
    with pytest.raises(ValueError):
        WaveFunctionUPS((3, 4), np.eye(12), gen, "tUPS", {"n_layers": 1})
    with pytest.raises(ValueError):
        WaveFunctionUPS((4, 4), np.eye(12), gen, "nonsense", {"n_layers": 1})
    # inactive/virtual bookkeeping and kappa indices
    wf2 = WaveFunctionUPS((2, 2), np.eye(12), gen, "fUCCSD", {})
    assert (wf2.num_inactive_orbs, wf2.num_active_orbs, wf2.num_virtual_orbs) == (5, 2, 5)
    assert len(wf2.kappa_idx) == 12 * 11 // 2 - (5 * 4 // 2) * 2 - 1


def test_term_adjoint_mask_form_matches_operator_loop():
    """The mask form of a term's adjoint used by the UCC kernels (sq_ucc.hip,
    part_fast / term_adjoint): for a spin part touching each orbital at most
    once, applying its operators in order (a creator needs its orbital
    occupied and clears it, an annihilator needs it empty and sets it, each
    adding popc(m & below(orbital)) to the parity) is the same as: applies iff
    m holds every creator orbital and no annihilator orbital; image m ^ s ^ c;
    parity popc(m & x) + c0, x the XOR of the below-masks, c0 the sum of
    popc(flipped-so-far & below).  Brute force over random strings."""
    import random

    def loop(ops, m):
        par = 0
        for orb, cre in ops:
            bit = 1 << orb
            par += bin(m & (bit - 1)).count("1")
            if cre:
                if not m & bit:
                    return None
                m &= ~bit
            else:
                if m & bit:
                    return None
                m |= bit
        return m, par

    def mask_form(ops):
        s = c = x = c0 = flipped = 0
        for orb, cre in ops:
            bit = 1 << orb
            x ^= bit - 1
            c0 += bin(flipped & (bit - 1)).count("1")
            flipped |= bit
            if cre:
                s |= bit
            else:
                c |= bit
        return s, c, x, c0

    rng = random.Random(7)
    applied = 0
    for _ in range(50000):
        orbs = rng.sample(range(14), rng.randint(0, 4))
        ops = [(o, rng.randint(0, 1)) for o in orbs]
        m = rng.getrandbits(14)
        ref = loop(ops, m)
        s, c, x, c0 = mask_form(ops)
        passes = (m & s) == s and not m & c
        assert passes == (ref is not None)
        if ref is None:
            continue
        applied += 1
        assert ref[0] == m ^ s ^ c
        assert ref[1] % 2 == (bin(m & x).count("1") + c0) % 2
    assert applied > 10000
