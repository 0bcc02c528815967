"""Second-quantised operators as dicts of operator strings.

Same representation and canonical ordering as the reference's
FermionicOperator (fermionic_operator.py:27-488): keys are tuples of
(spin_orbital, is_creator); products are brought to normal order with
creators first in descending index, then annihilators in descending index,
contractions a_p a+_p = 1 - a+_p a_p expanded (a fully contracted empty term is
dropped, as in the reference), and merged coefficients with |c| < 1e-14 are
removed.  ``get_folded_operator`` restricts an operator to the active space
with doubly occupied inactive and empty virtual orbitals.

These objects describe WHAT the engine applies; the application itself runs
on the GPU (operator_state_algebra.propagate_state).
"""

from __future__ import annotations

from typing import Iterable

_DROP = 1e-14


def _merge(dst: dict, key, c) -> None:
    if key in dst:
        dst[key] += c
        if abs(dst[key]) < _DROP:
            del dst[key]
    else:
        dst[key] = c


def _canonical(ops: list, c: float, out: dict) -> None:
    """Normal-order one product string into ``out`` (adjacent transpositions)."""
    work = [(ops, c)]
    while work:
        cur, f = work.pop()
        cur = list(cur)
        dead = False
        moved = True
        while moved and not dead:
            moved = False
            for i in range(len(cur) - 1):
                (p, cp), (q, cq) = cur[i], cur[i + 1]
                if cp == cq:
                    # both creators or both annihilators: descending index
                    if p == q:
                        dead = True
                        break
                    if p < q:
                        cur[i], cur[i + 1] = cur[i + 1], cur[i]
                        f = -f
                        moved = True
                elif cq and not cp:
                    # annihilator left of a creator
                    if p == q:
                        rest = cur[:i] + cur[i + 2 :]
                        if rest:
                            work.append((rest, f))
                    cur[i], cur[i + 1] = cur[i + 1], cur[i]
                    f = -f
                    moved = True
        if not dead:
            _merge(out, tuple(cur), f)


def do_extended_normal_ordering(fermistring: "FermionicOperator") -> dict:
    out: dict = {}
    for key, c in fermistring.operators.items():
        _canonical(list(key), c, out)
    return out


class FermionicOperator:
    __slots__ = ("operators",)

    def __init__(self, annihilation_operator: dict) -> None:
        if not isinstance(annihilation_operator, dict):
            raise ValueError(f"Could not assign operator of {type(annihilation_operator)}.")
        self.operators = annihilation_operator

    def __add__(self, other: "FermionicOperator") -> "FermionicOperator":
        ops = dict(self.operators)
        for k, v in other.operators.items():
            _merge(ops, k, v)
        return FermionicOperator(ops)

    def __iadd__(self, other: "FermionicOperator") -> "FermionicOperator":
        for k, v in other.operators.items():
            _merge(self.operators, k, v)
        return self

    def __sub__(self, other: "FermionicOperator") -> "FermionicOperator":
        ops = dict(self.operators)
        for k, v in other.operators.items():
            _merge(ops, k, -v)
        return FermionicOperator(ops)

    def __isub__(self, other: "FermionicOperator") -> "FermionicOperator":
        for k, v in other.operators.items():
            _merge(self.operators, k, -v)
        return self

    def _product(self, other: "FermionicOperator") -> dict:
        ops: dict = {}
        for k1, v1 in other.operators.items():
            for k2, v2 in self.operators.items():
                part: dict = {}
                _canonical(list(k2 + k1), v2 * v1, part)
                for k, v in part.items():
                    _merge(ops, k, v)
        return ops

    def __mul__(self, other):
        if type(other) in (float, int):
            return FermionicOperator({k: v * other for k, v in self.operators.items()})
        if isinstance(other, FermionicOperator):
            return FermionicOperator(self._product(other))
        raise TypeError(f"Got unknown type of fermistring: {type(other)}")

    def __imul__(self, other):
        if type(other) in (float, int):
            for k in self.operators:
                self.operators[k] *= other
            return self
        if isinstance(other, FermionicOperator):
            self.operators = self._product(other)
            return self
        raise TypeError(f"Got unknown type of fermistring: {type(other)}")

    def __rmul__(self, number: float) -> "FermionicOperator":
        return FermionicOperator({k: v * number for k, v in self.operators.items()})

    def __neg__(self) -> "FermionicOperator":
        return FermionicOperator({k: -v for k, v in self.operators.items()})

    @property
    def dagger(self) -> "FermionicOperator":
        flipped = {tuple((i, not d) for i, d in reversed(k)): v for k, v in self.operators.items()}
        return FermionicOperator(do_extended_normal_ordering(FermionicOperator(flipped)))

    @property
    def operator_count(self) -> dict[int, int]:
        cnt: dict[int, int] = {}
        for k in self.operators:
            cnt[len(k)] = cnt.get(len(k), 0) + 1
        return cnt

    @property
    def operators_readable(self) -> dict[str, float]:
        return {"".join(("c" if d else "a") + str(i) for i, d in k): v for k, v in self.operators.items()}

    def get_folded_operator(self, num_inactive_orbs: int, num_active_orbs: int, num_virtual_orbs: int):
        """Active-space part with inactive orbitals doubly occupied and virtuals empty.

        Same rules as fermionic_operator.py:379-471: strings touching a virtual
        orbital vanish; the inactive creators must equal the inactive
        annihilators (as ordered lists); a sign comes from moving the inactive
        part past the active creators and from reversing the ket side.
        Active indices are shifted so the first active spin orbital is 0.
        """
        lo = 2 * num_inactive_orbs
        hi = lo + 2 * num_active_orbs
        out: dict = {}
        for key, coeff in self.operators.items():
            in_c, in_a, act_c, act_a = [], [], [], []
            skip = False
            for idx, dag in key:
                if idx >= hi:
                    skip = True
                    break
                if idx < lo:
                    (in_c if dag else in_a).append(idx)
                elif dag:
                    act_c.append((idx - lo, True))
                else:
                    act_a.append((idx - lo, False))
            if skip or in_c != in_a:
                continue
            sign = -1 if (len(in_c) % 2 == 1 and len(act_c) % 2 == 1) else 1
            if (len(in_a) // 2) % 2 == 1:
                sign = -sign
            nk = tuple(act_c + act_a)
            out[nk] = out.get(nk, 0.0) + sign * coeff
        return FermionicOperator(out)


def terms_of(ops: Iterable) -> list:
    """Flatten FermionicOperator keys into the engine's string format."""
    return [(k, v) for op in ops for k, v in op.operators.items()]
