"""Ansatz layouts: theta index -> (excitation type, indices).

Mirrors the gate lists of the reference's UpsStructure / UccStructure
(util.py:547-1073) and their excitation iterators (util.py:5-544): the order
of gates, their index tuples, grad_param_R and param_names must match the
reference exactly because theta vectors are exchanged between the two codes.
"""

from __future__ import annotations

import itertools
from collections.abc import Generator, Sequence
from typing import Any


def _spin_ok(creators: Sequence[int], annihilators: Sequence[int]) -> bool:
    """True when the excitation conserves the alpha and beta electron counts."""
    na = sum(1 for x in creators if x % 2 == 0) - sum(1 for x in annihilators if x % 2 == 0)
    nb = sum(1 for x in creators if x % 2 == 1) - sum(1 for x in annihilators if x % 2 == 1)
    return na == 0 and nb == 0


def iterate_t1_sa(occ: Sequence[int], unocc: Sequence[int]) -> Generator[tuple[int, int, float], None, None]:
    """Spin-adapted singles (util.py:5-21): yields (a, i, 1/sqrt 2)."""
    for i in occ:
        for a in unocc:
            yield a, i, 2 ** (-1 / 2)


def _sa_double_case(i: int, j: int, a: int, b: int) -> int:
    if i == j and a == b:
        return 1
    if i == j:
        return 2
    if a == b:
        return 3
    return 4


def _t2_sa_from(pairs_ij, pairs_ab) -> Generator[tuple[int, int, int, int, float, int], None, None]:
    for i, j in pairs_ij:
        for a, b in pairs_ab(i, j):
            fac = (2.0 if a == b else 1.0) * (2.0 if i == j else 1.0)
            yield a, i, b, j, 0.5 * fac ** (-1 / 2), _sa_double_case(i, j, a, b)
            if i != j and a != b:
                yield a, i, b, j, 1 / (2 * 3 ** (1 / 2)), 5


def iterate_t2_sa(occ: Sequence[int], unocc: Sequence[int]):
    """Spin-adapted doubles (util.py:24-58): yields (a, i, b, j, fac, case)."""
    pij = [(occ[x], occ[y]) for x in range(len(occ)) for y in range(x, len(occ))]
    pab = [(unocc[x], unocc[y]) for x in range(len(unocc)) for y in range(x, len(unocc))]
    return _t2_sa_from(pij, lambda i, j: pab)


def iterate_t1_sa_generalized(num_orbs: int):
    """Generalized spin-adapted singles (util.py:61-77)."""
    for i in range(num_orbs):
        for a in range(i + 1, num_orbs):
            yield a, i, 2 ** (-1 / 2)


def iterate_t2_sa_generalized(num_orbs: int):
    """Generalized spin-adapted doubles (util.py:80-110)."""
    pij = [(i, j) for i in range(num_orbs) for j in range(i, num_orbs)]

    def pab(i, j):
        return [(a, b) for a in range(max(i, j) + 1, num_orbs) for b in range(a, num_orbs)]

    return _t2_sa_from(pij, pab)


def iterate_t1(occ_spin: Sequence[int], unocc_spin: Sequence[int]):
    """Spin-conserving singles (util.py:113-137): yields (a, i)."""
    for a in unocc_spin:
        for i in occ_spin:
            if _spin_ok([a], [i]):
                yield a, i


def iterate_t1_generalized(num_spin_orbs: int):
    """util.py:140-164."""
    for i in range(num_spin_orbs):
        for a in range(i + 1, num_spin_orbs):
            if _spin_ok([a], [i]):
                yield a, i


def _iterate_tn(occ_spin: Sequence[int], unocc_spin: Sequence[int], rank: int):
    """Spin-conserving rank-n excitations in the reference's nesting order.

    The reference nests the creator loops outside the annihilator loops, each
    over strictly increasing positions (util.py:167-510), and yields the
    interleaved tuple (a, i, b, j, ...).
    """
    for cre in itertools.combinations(range(len(unocc_spin)), rank):
        cs = [unocc_spin[x] for x in cre]
        for ann in itertools.combinations(range(len(occ_spin)), rank):
            an = [occ_spin[x] for x in ann]
            if _spin_ok(cs, an):
                out = []
                for c, a in zip(cs, an):
                    out += [c, a]
                yield tuple(out)


def iterate_t2(occ_spin, unocc_spin):
    return _iterate_tn(occ_spin, unocc_spin, 2)


def iterate_t3(occ_spin, unocc_spin):
    return _iterate_tn(occ_spin, unocc_spin, 3)


def iterate_t4(occ_spin, unocc_spin):
    return _iterate_tn(occ_spin, unocc_spin, 4)


def iterate_t5(occ_spin, unocc_spin):
    return _iterate_tn(occ_spin, unocc_spin, 5)


def iterate_t6(occ_spin, unocc_spin):
    return _iterate_tn(occ_spin, unocc_spin, 6)


def iterate_t2_generalized(num_spin_orbs: int):
    """util.py:193-247: i <= j < a <= b ordering, spin conserving."""
    for i in range(num_spin_orbs):
        for j in range(i, num_spin_orbs):
            for a in range(max(i, j) + 1, num_spin_orbs):
                for b in range(a, num_spin_orbs):
                    if _spin_ok([a, b], [i, j]):
                        yield a, i, b, j


def iterate_pair_t2(occ: Sequence[int], unocc: Sequence[int]):
    """Pair doubles (util.py:513-527)."""
    for i in occ:
        for a in unocc:
            yield 2 * a, 2 * i, 2 * a + 1, 2 * i + 1


def iterate_pair_t2_generalized(num_orbs: int):
    """util.py:530-544."""
    for i in range(num_orbs):
        for a in range(i + 1, num_orbs):
            yield 2 * a, 2 * i, 2 * a + 1, 2 * i + 1


class UccStructure:
    """Unfactorised UCC excitation list (util.py:547-639)."""

    __slots__ = ("excitation_indices", "excitation_operator_type", "n_params")

    def __init__(self) -> None:
        self.excitation_indices: list[tuple[int, ...]] = []
        self.excitation_operator_type: list[str] = []
        self.n_params = 0

    def _add(self, kind: str, idx: tuple[int, ...]) -> None:
        self.excitation_indices.append(idx)
        self.excitation_operator_type.append(kind)
        self.n_params += 1

    def add_sa_singles(self, occ, unocc) -> None:
        for a, i, _ in iterate_t1_sa(occ, unocc):
            self._add("sa_single", (i, a))

    def add_sa_doubles(self, occ, unocc) -> None:
        for a, i, b, j, _, case in iterate_t2_sa(occ, unocc):
            self._add(f"sa_double_{case}", (i, j, a, b))

    def _add_rank(self, occ_spin, unocc_spin, rank, kind):
        for t in _iterate_tn(occ_spin, unocc_spin, rank):
            self._add(kind, tuple(t[1::2]) + tuple(t[0::2]))

    def add_triples(self, occ_spin, unocc_spin) -> None:
        self._add_rank(occ_spin, unocc_spin, 3, "triple")

    def add_quadruples(self, occ_spin, unocc_spin) -> None:
        self._add_rank(occ_spin, unocc_spin, 4, "quadruple")

    def add_quintuples(self, occ_spin, unocc_spin) -> None:
        self._add_rank(occ_spin, unocc_spin, 5, "quintuple")

    def add_sextuples(self, occ_spin, unocc_spin) -> None:
        self._add_rank(occ_spin, unocc_spin, 6, "sextuple")


class UpsStructure:
    """Unitary product state gate list (util.py:642-1073)."""

    __slots__ = ("excitation_indices", "excitation_operator_type", "grad_param_R", "n_params", "param_names")

    def __init__(self) -> None:
        self.excitation_indices: list[tuple[int, ...]] = []
        self.excitation_operator_type: list[str] = []
        self.n_params: int = 0
        self.grad_param_R: dict[str, int] = {}
        self.param_names: list[str] = []

    def _add(self, kind: str, idx: tuple[int, ...], R: int | None) -> None:
        name = f"p{self.n_params:09d}"
        self.excitation_operator_type.append(kind)
        self.excitation_indices.append(tuple(idx))
        if R is not None:
            self.grad_param_R[name] = R
        self.param_names.append(name)
        self.n_params += 1

    def create_tiled(self, num_active_orbs: int, ansatz_options: dict[str, Any]) -> None:
        """tUPS / QNP brick wall (util.py:653-745)."""
        valid = ("n_layers", "do_qnp", "skip_last_singles", "do_tups", "do_pp")
        for opt in ansatz_options:
            if opt not in valid:
                raise ValueError(f"Got unknown option for tUPS, {opt}. Valid options are: {valid}")
        if "n_layers" not in ansatz_options:
            raise ValueError("tUPS require the option 'n_layers'")
        do_tups = ansatz_options.get("do_tups", False)
        do_qnp = ansatz_options.get("do_qnp", False)
        if not do_tups and not do_qnp:
            raise ValueError("No tiled ansatz specified.")
        if do_tups and do_qnp:
            raise ValueError("More than one tiled ansatz specfied.")
        skip_last = ansatz_options.get("skip_last_singles", False)
        n_layers = ansatz_options["n_layers"]
        n = num_active_orbs
        for layer in range(n_layers):
            last = layer + 1 == n_layers
            for start in (0, 1):
                for p in range(start, n - 1, 2):
                    if do_tups:
                        self._add("sa_single", (p, p + 1), 4)
                    self._add("double", (2 * p, 2 * p + 1, 2 * p + 2, 2 * p + 3), 2)
                    if start == 0 and last and skip_last and n == 2:
                        continue
                    if start == 1 and last and skip_last:
                        continue
                    self._add("sa_single", (p, p + 1), 4)

    def create_fUCC(self, occ_idx, unocc_idx, occ_spin_idx, unocc_spin_idx, num_orbs, ansatz_options) -> None:
        """Factorised UCC (util.py:747-953)."""
        valid = ("n_layers", "S", "D", "SAGS", "pD", "GpD", "SAS", "T", "Q", "5", "6", "SAD", "GS", "GD")
        for opt in ansatz_options:
            if opt not in valid:
                raise ValueError(f"Got unknown option for fUCC, {opt}. Valid options are: {valid}")
        if "n_layers" not in ansatz_options:
            raise ValueError("fUCC require the option 'n_layers'")
        flags = {k: bool(ansatz_options.get(k, False)) for k in valid[1:]}
        if not any(flags.values()):
            raise ValueError("fUCC requires some excitations got none.")
        for _ in range(ansatz_options["n_layers"]):
            if flags["S"]:
                for a, i in iterate_t1(occ_spin_idx, unocc_spin_idx):
                    self._add("single", (i, a), 2)
            if flags["GS"]:
                for a, i in iterate_t1_generalized(2 * num_orbs):
                    self._add("single", (i, a), 2)
            if flags["SAS"]:
                for a, i, _f in iterate_t1_sa(occ_idx, unocc_idx):
                    self._add("sa_single", (i, a), 4)
            if flags["SAGS"]:
                for a, i, _f in iterate_t1_sa_generalized(num_orbs):
                    self._add("sa_single", (i, a), 4)
            if flags["D"]:
                for a, i, b, j in iterate_t2(occ_spin_idx, unocc_spin_idx):
                    self._add("double", (i, j, a, b), 2)
            if flags["GD"]:
                for a, i, b, j in iterate_t2_generalized(2 * num_orbs):
                    self._add("double", (i, j, a, b), 2)
            if flags["pD"]:
                for a, i, b, j in iterate_pair_t2(occ_idx, unocc_idx):
                    self._add("double", (i, j, a, b), 2)
            if flags["GpD"]:
                for a, i, b, j in iterate_pair_t2_generalized(num_orbs):
                    self._add("double", (i, j, a, b), 2)
            for key, rank, kind in (("T", 3, "triple"), ("Q", 4, "quadruple"), ("5", 5, "quintuple"), ("6", 6, "sextuple")):
                if flags[key]:
                    for t in _iterate_tn(occ_spin_idx, unocc_spin_idx, rank):
                        self._add(kind, tuple(t[1::2]) + tuple(t[0::2]), 2)
            if flags["SAD"]:
                for a, i, b, j, _f, case in iterate_t2_sa(occ_idx, unocc_idx):
                    # The reference gives SA doubles no RotoSolve period.
                    self._add(f"sa_double_{case}", (i, j, a, b), None)

    def create_SDSfUCC(self, occ_idx, unocc_idx, occ_spin_idx, unocc_spin_idx, num_orbs, ansatz_options) -> None:
        """SDS-ordered factorised UCC (util.py:955-1073)."""
        valid = ("n_layers", "D", "pD", "GpD")
        for opt in ansatz_options:
            if opt not in valid:
                raise ValueError(f"Got unknown option for SDSfUCC, {opt}. Valid options are: {valid}")
        if "n_layers" not in ansatz_options:
            raise ValueError("SDSfUCC require the option 'n_layers'")
        do_D = ansatz_options.get("D", False)
        do_pD = ansatz_options.get("pD", False)
        do_GpD = ansatz_options.get("GpD", False)
        if not (do_D or do_pD or do_GpD):
            raise ValueError("SDSfUCC requires some excitations got none.")
        for _ in range(ansatz_options["n_layers"]):
            if do_D:
                for a, i, b, j in iterate_t2(occ_spin_idx, unocc_spin_idx):
                    same = i % 2 == a % 2
                    self._add("single", (i, a) if same else (i, b), 2)
                    self._add("double", (i, j, a, b), 2)
                    self._add("single", (j, b) if same else (j, a), 2)
            if do_pD:
                # The reference emits a "double" with a two-index tuple here
                # (util.py:1043-1048); kept as is so layouts match.
                for a, i, b, j in iterate_pair_t2(occ_idx, unocc_idx):
                    self._add("double", (i // 2, a // 2), 4)
                    self._add("double", (i, j, a, b), 2)
                    self._add("sa_single", (i // 2, a // 2), 4)
            if do_GpD:
                for a, i, b, j in iterate_pair_t2_generalized(num_orbs):
                    self._add("sa_single", (i // 2, a // 2), 4)
                    self._add("double", (i, j, a, b), 2)
                    self._add("sa_single", (i // 2, a // 2), 4)
