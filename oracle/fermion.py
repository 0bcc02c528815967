"""Fermionic operator algebra restated for the oracle (test infrastructure only).

Operators are dicts {((spin_orbital, is_creator), ...): coefficient}, the
reference's representation (fermionic_operator.py:107-126).  Products are put in
the reference's canonical order by ``normal_order``: creators first with
descending index, then annihilators with descending index, contractions
a_p a+_p -> delta - a+_p a_p expanded (fermionic_operator.py:27-104).
Coefficients with |c| < 1e-14 are dropped when terms merge, as in the
reference's add/sub/mul (fermionic_operator.py:127-239).
"""

from __future__ import annotations

import math

Key = tuple  # tuple[tuple[int, bool], ...]


def _order_one(ops: list, coeff: float, out: dict) -> None:
    """Normal-order one product string, accumulating canonical terms into out."""
    stack = [(list(ops), coeff)]
    while stack:
        cur, c = stack.pop()
        changed = True
        zero = False
        while changed and not zero:
            changed = False
            for i in range(len(cur) - 1):
                (p, dp), (q, dq) = cur[i], cur[i + 1]
                if dp and dq:
                    if p == q:
                        zero = True
                        break
                    if p < q:
                        cur[i], cur[i + 1] = cur[i + 1], cur[i]
                        c = -c
                        changed = True
                elif (not dp) and dq:
                    if p == q:
                        contracted = cur[:i] + cur[i + 2 :]
                        # A fully contracted (empty) string is dropped, as the
                        # reference does (fermionic_operator.py:69-72).
                        if contracted:
                            stack.append((contracted, c))
                    cur[i], cur[i + 1] = cur[i + 1], cur[i]
                    c = -c
                    changed = True
                elif dp and not dq:
                    pass
                else:
                    if p == q:
                        zero = True
                        break
                    if p < q:
                        cur[i], cur[i + 1] = cur[i + 1], cur[i]
                        c = -c
                        changed = True
        if zero:
            continue
        key = tuple(cur)
        if key in out:
            out[key] += c
            if abs(out[key]) < 1e-14:
                del out[key]
        else:
            out[key] = c


def normal_order(terms: dict) -> dict:
    out: dict = {}
    for key, c in terms.items():
        _order_one(list(key), c, out)
    return out


class FOp:
    """Minimal restatement of FermionicOperator (fermionic_operator.py:107-321)."""

    __slots__ = ("operators",)

    def __init__(self, operators: dict) -> None:
        self.operators = operators

    def __add__(self, other: "FOp") -> "FOp":
        ops = dict(self.operators)
        for k, v in other.operators.items():
            if k in ops:
                ops[k] += v
                if abs(ops[k]) < 1e-14:
                    del ops[k]
            else:
                ops[k] = v
        return FOp(ops)

    def __neg__(self) -> "FOp":
        return FOp({k: -v for k, v in self.operators.items()})

    def __sub__(self, other: "FOp") -> "FOp":
        return self + (-other)

    def __mul__(self, other):
        if isinstance(other, (int, float)):
            return FOp({k: v * other for k, v in self.operators.items()})
        ops: dict = {}
        for k1, v1 in other.operators.items():
            for k2, v2 in self.operators.items():
                for k, v in normal_order({k2 + k1: v2 * v1}).items():
                    if k in ops:
                        ops[k] += v
                        if abs(ops[k]) < 1e-14:
                            del ops[k]
                    else:
                        ops[k] = v
        return FOp(ops)

    def __rmul__(self, number: float) -> "FOp":
        return FOp({k: v * number for k, v in self.operators.items()})

    @property
    def dagger(self) -> "FOp":
        ops = {}
        for k, v in self.operators.items():
            ops[tuple((i, not d) for i, d in reversed(k))] = v
        return FOp(normal_order(ops))

    def folded(self, n_inact: int, n_act: int, n_virt: int) -> "FOp":
        """Restates get_folded_operator (fermionic_operator.py:379-471)."""
        out: dict = {}
        lo, hi = 2 * n_inact, 2 * n_inact + 2 * n_act
        for key, coeff in self.operators.items():
            ina_d, ina, act_d, act, virt = [], [], [], [], False
            for idx, dag in key:
                if idx >= hi:
                    virt = True
                    break
                if idx < lo:
                    (ina_d if dag else ina).append(idx)
                else:
                    (act_d if dag else act).append((idx - lo, dag))
            if virt or ina_d != ina:
                continue
            fac = 1
            if len(ina_d) % 2 == 1 and len(act_d) % 2 == 1:
                fac = -fac
            for i in range(1, len(ina) + 1):
                if i % 2 == 0:
                    fac = -fac
            nk = tuple(act_d + act)
            out[nk] = out.get(nk, 0.0) + fac * coeff
        return FOp(out)


def a_spin(idx: int, dagger: bool) -> FOp:
    return FOp({((idx, dagger),): 1})


def Epq(p: int, q: int) -> FOp:
    """E_pq = a+_pa a_qa + a+_pb a_qb (operators.py:40-58)."""
    return a_spin(2 * p, True) * a_spin(2 * q, False) + a_spin(2 * p + 1, True) * a_spin(2 * q + 1, False)


def epqrs(p: int, q: int, r: int, s: int) -> FOp:
    """e_pqrs = E_pq E_rs - delta_qr E_ps (operators.py:61-80)."""
    op = Epq(p, q) * Epq(r, s)
    if q == r:
        op = op - Epq(p, s)
    return op


def _gen(creators, annihilators, anti_hermitian):
    op = a_spin(creators[0], True)
    for c in creators[1:]:
        op = op * a_spin(c, True)
    for a in annihilators:
        op = op * a_spin(a, False)
    if anti_hermitian:
        op = op - op.dagger
    return op


def G1(i, a, ah=False):
    """a+_a a_i (operators.py:145-163)."""
    return _gen([a], [i], ah)


def G2(i, j, a, b, ah=False):
    """a+_a a+_b a_j a_i (operators.py:166-188)."""
    return _gen([a, b], [j, i], ah)


def G3(i, j, k, a, b, c, ah=False):
    return _gen([a, b, c], [k, j, i], ah)


def G4(i, j, k, l, a, b, c, d, ah=False):
    return _gen([a, b, c, d], [l, k, j, i], ah)


def G5(i, j, k, l, m, a, b, c, d, e, ah=False):
    return _gen([a, b, c, d, e], [m, l, k, j, i], ah)


def G6(i, j, k, l, m, n, a, b, c, d, e, f, ah=False):
    return _gen([a, b, c, d, e, f], [n, m, l, k, j, i], ah)


def G1_sa(i, a, ah=False):
    """2^-1/2 E_ai (operators.py:362-379)."""
    op = 2 ** (-1 / 2) * Epq(a, i)
    if ah:
        op = op - op.dagger
    return op


def G2_sa(i, j, a, b, case, ah=False):
    """Spin-adapted doubles, cases 1-5 (operators.py:382-443)."""
    if case in (1, 2, 3, 4):
        fac = 1
        if a == b:
            fac *= 2
        if i == j:
            fac *= 2
        op = 1 / 2 * fac ** (-1 / 2) * (Epq(a, i) * Epq(b, j) + Epq(a, j) * Epq(b, i))
    elif case == 5:
        op = 1 / (2 * 3 ** (1 / 2)) * (Epq(a, i) * Epq(b, j) - Epq(a, j) * Epq(b, i))
    else:
        raise ValueError(case)
    if ah:
        op = op - op.dagger
    return op


def hamiltonian_0i_0a(h, g, n_inact: int, n_act: int) -> FOp:
    """Restates hamiltonian_0i_0a (operators.py:476-529) incl. its 1e-14 screens."""
    H = FOp({})
    act = range(n_inact, n_inact + n_act)
    for i in range(n_inact):
        if abs(h[i, i]) > 1e-14:
            H = H + h[i, i] * Epq(i, i)
    for p in act:
        for q in act:
            if abs(h[p, q]) > 1e-14:
                H = H + h[p, q] * Epq(p, q)
    for i in range(n_inact):
        for j in range(n_inact):
            if abs(g[i, i, j, j]) > 1e-14:
                H = H + 1 / 2 * g[i, i, j, j] * epqrs(i, i, j, j)
            if i != j and abs(g[j, i, i, j]) > 1e-14:
                H = H + 1 / 2 * g[j, i, i, j] * epqrs(j, i, i, j)
    for i in range(n_inact):
        for p in act:
            for q in act:
                if abs(g[i, i, p, q]) > 1e-14:
                    H = H + 1 / 2 * g[i, i, p, q] * epqrs(i, i, p, q)
                if abs(g[p, q, i, i]) > 1e-14:
                    H = H + 1 / 2 * g[p, q, i, i] * epqrs(p, q, i, i)
                if abs(g[p, i, i, q]) > 1e-14:
                    H = H + 1 / 2 * g[p, i, i, q] * epqrs(p, i, i, q)
                if abs(g[i, p, q, i]) > 1e-14:
                    H = H + 1 / 2 * g[i, p, q, i] * epqrs(i, p, q, i)
    for p in act:
        for q in act:
            for r in act:
                for s in act:
                    if abs(g[p, q, r, s]) > 1e-14:
                        H = H + 1 / 2 * g[p, q, r, s] * epqrs(p, q, r, s)
    return H


def hamiltonian_active_fast(h, g, n_act: int) -> FOp:
    """Normal-ordered active-space H for the no-inactive case, assembled in
    closed form instead of through n^4 operator products.

    Same operator as hamiltonian_0i_0a with n_inact = 0 (checked in the oracle
    tests); used only to make large-n CPU baselines affordable.  The terms are
    e_pqrs = sum_{st} a+_{p s} a+_{r t} a_{s t} a_{q s} (with operator order
    fixed by normal_order), plus the one-body part.
    """
    out: dict = {}

    def add(key_ops, c):
        for k, v in normal_order({tuple(key_ops): c}).items():
            if k in out:
                out[k] += v
                if abs(out[k]) < 1e-14:
                    del out[k]
            else:
                out[k] = v

    for p in range(n_act):
        for q in range(n_act):
            if abs(h[p, q]) > 1e-14:
                for s in (0, 1):
                    add([(2 * p + s, True), (2 * q + s, False)], h[p, q])
    for p in range(n_act):
        for q in range(n_act):
            for r in range(n_act):
                for s_ in range(n_act):
                    c = g[p, q, r, s_]
                    if abs(c) <= 1e-14:
                        continue
                    for sa in (0, 1):
                        for sb in (0, 1):
                            # E_pq E_rs - delta_qr E_ps == a+_p a+_r a_s a_q (spin-summed)
                            add(
                                [(2 * p + sa, True), (2 * r + sb, True), (2 * s_ + sb, False), (2 * q + sa, False)],
                                0.5 * c,
                            )
    return FOp(out)


def sa_double_tables():
    """Frequencies S and T^k coefficient tables of the multi-frequency SA doubles.

    Restated from operator_state_algebra.py:1086-1279 (construct_ups_state).
    Odd powers multiply sin(S_f theta), even powers multiply (cos(S_f theta) - 1).
    Returns {case: (S, [k1, k2, ...])}.
    """
    r2, r3 = math.sqrt(2), math.sqrt(3)
    t23 = (
        (1, r2 / 2),
        [(-1, 2 * r2), (1, -4), (-2, 2 * r2), (2, -4)],
    )
    t4 = (
        (1, r2, r2 / 2, 1 / 2),
        [
            (2 / 3, -r2 / 42, -8 * r2 / 3, 128 / 21),
            (-2 / 3, 1 / 42, 16 / 3, -256 / 21),
            (13 / 3, -r2 / 6, -44 * r2 / 3, 64 / 3),
            (-13 / 3, 1 / 6, 88 / 3, -128 / 3),
            (22 / 3, -r2 / 3, -52 * r2 / 3, 64 / 3),
            (-22 / 3, 1 / 3, 104 / 3, -128 / 3),
            (8 / 3, -4 * r2 / 21, -16 * r2 / 3, 128 / 21),
            (-8 / 3, 4 / 21, 32 / 3, -256 / 21),
        ],
    )
    t5 = (
        (r2, r2 / 2, r3 / 3, r3 / 2, r3 / 6),
        [
            (r2 / 1150, 8 * r2 / 5, -54 * r3 / 25, -16 * r3 / 75, 432 * r3 / 115),
            (-1 / 1150, -16 / 5, 162 / 25, 32 / 75, -2592 / 115),
            (11 * r2 / 690, 404 * r2 / 15, -171 * r3 / 5, -56 * r3 / 15, 2952 * r3 / 115),
            (-11 / 690, -808 / 15, 513 / 5, 112 / 15, -17712 / 115),
            (133 * r2 / 1725, 308 * r2 / 3, -2718 * r3 / 25, -1192 * r3 / 75, 1368 * r3 / 23),
            (-133 / 1725, -616 / 3, 8154 / 25, 2384 / 75, -8208 / 23),
            (16 * r2 / 115, 608 * r2 / 5, -576 * r3 / 5, -112 * r3 / 5, 6192 * r3 / 115),
            (-16 / 115, -1216 / 5, 1728 / 5, 224 / 5, -37152 / 115),
            (48 * r2 / 575, 192 * r2 / 5, -864 * r3 / 25, -192 * r3 / 25, 1728 * r3 / 115),
            (-48 / 575, -384 / 5, 2592 / 25, 384 / 25, -10368 / 115),
        ],
    )
    return {2: t23, 3: t23, 4: t4, 5: t5}
