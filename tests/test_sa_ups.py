"""Host-side checks of the state-averaged UPS front end and of the dipole
integrals its oscillator strengths use (no GPU: construction touches no
device state)."""

from __future__ import annotations

import numpy as np
import pytest


def _h2():
    from slowquant_amd.hchain import HydrogenSlowQuant

    sq = HydrogenSlowQuant()
    sq.set_molecule("H 0.0 0.0 0.0; H 0.0 0.0 0.735", distance_unit="angstrom")
    sq.set_basis_set("STO-3G")
    sq.init_hartree_fock()
    sq.hartree_fock.run_restricted_hartree_fock()
    return sq


H2_STATES = ([[1], [2 ** (-1 / 2), -(2 ** (-1 / 2))], [1]], [["1100"], ["1001", "0110"], ["0011"]])


def test_sa_constructor_places_reference_states():
    from slowquant_amd.sa_ups_wavefunction import WaveFunctionSAUPS

    sq = _h2()
    wf = WaveFunctionSAUPS((2, 2), sq.hartree_fock.mo_coeff, sq, H2_STATES, "tUPS",
                           {"n_layers": 1, "skip_last_singles": True}, include_active_kappa=True)
    assert wf.num_states == 3
    assert wf.csf_coeffs.shape == (3, 4)
    d2i = wf.ci_info.det2idx
    assert wf.csf_coeffs[0, d2i[int("1100", 2)]] == 1.0
    assert wf.csf_coeffs[1, d2i[int("1001", 2)]] == pytest.approx(2 ** -0.5)
    assert wf.csf_coeffs[1, d2i[int("0110", 2)]] == pytest.approx(-(2 ** -0.5))
    assert wf.csf_coeffs[2, d2i[int("0011", 2)]] == 1.0
    # before any theta is set the states are the references themselves
    assert np.array_equal(wf.ci_coeffs, wf.csf_coeffs)
    assert len(wf.thetas) == wf.ups_layout.n_params
    with pytest.raises(ValueError, match="Expected"):
        wf.thetas = [0.0] * (len(wf.thetas) + 1)


@pytest.mark.parametrize(
    "states,msg",
    [
        (([[1, 0.5]], [["1100"]]), "Mismatch in number of coefficients"),
        (([[1]], [["110"]]), "Length of determinant"),
        (([[0.9]], [["1100"]]), "is not normalized"),
        (([[1], [2 ** -0.5, 2 ** -0.5]], [["1100"], ["1100", "0011"]]), "are not orthogonal"),
    ],
)
def test_sa_constructor_rejects_bad_states(states, msg):
    from slowquant_amd.sa_ups_wavefunction import WaveFunctionSAUPS

    sq = _h2()
    with pytest.raises(ValueError, match=msg):
        WaveFunctionSAUPS((2, 2), sq.hartree_fock.mo_coeff, sq, states, "tUPS", {"n_layers": 1})


def test_sa_constructor_rejects_pp_and_unknown_ansatz():
    from slowquant_amd.sa_ups_wavefunction import WaveFunctionSAUPS

    sq = _h2()
    with pytest.raises(ValueError, match="perfect pairing"):
        WaveFunctionSAUPS((2, 2), sq.hartree_fock.mo_coeff, sq, H2_STATES, "tUPS", {"n_layers": 1, "do_pp": True})
    with pytest.raises(ValueError, match="unknown ansatz"):
        WaveFunctionSAUPS((2, 2), sq.hartree_fock.mo_coeff, sq, H2_STATES, "fUCCSD", {})


def _dipole_by_quadrature(ints, coords, k):
    """<a|r_k|b> as a sum over primitive pairs of products of three 1-D
    integrals, each by a dense trapezoid rule (independent of the Gaussian
    product theorem used by the closed form)."""
    x = np.linspace(-12.0, 12.0, 24001)
    n = len(coords)
    M = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            for a in range(3):
                for b in range(3):
                    val = ints._d[a] * ints._d[b]
                    for dim in range(3):
                        f = np.exp(-ints._alpha[a] * (x - coords[i, dim]) ** 2 - ints._alpha[b] * (x - coords[j, dim]) ** 2)
                        if dim == k:
                            f = f * x
                        val *= np.trapezoid(f, x)
                    M[i, j] += val
    return M


def test_dipole_integrals_against_quadrature_and_translation():
    from slowquant_amd.hchain import HydrogenSlowQuant

    sq = HydrogenSlowQuant()
    sq.set_molecule("H -0.45 -0.3897114317 0.1; H 0.45 -0.3897114317 0.0; H 0.0 0.3897114317 -0.2",
                    distance_unit="angstrom", molecular_charge=1)
    sq.set_basis_set("STO-3G")
    ints = sq.integral
    coords = sq.molecule.coordinates
    S = ints.get_multipole_matrix(np.array([0, 0, 0]))
    assert np.allclose(np.diag(S), 1.0)
    for k in range(3):
        carts = np.zeros(3, dtype=int)
        carts[k] = 1
        D = ints.get_multipole_matrix(carts)
        assert np.abs(D - D.T).max() < 1e-14
        # a normalised 1s function's centroid is its centre
        assert np.abs(np.diag(D) - coords[:, k]).max() < 1e-12
        assert np.abs(D - _dipole_by_quadrature(ints, coords, k)).max() < 1e-9
        # shifting every atom by t adds t_k S
        t = np.array([0.3, -0.7, 1.1])
        sq2 = HydrogenSlowQuant()
        sq2.set_molecule(";".join(f"H {c[0]} {c[1]} {c[2]}" for c in coords + t), distance_unit="bohr",
                         molecular_charge=1)
        sq2.set_basis_set("STO-3G")
        assert np.abs(sq2.integral.get_multipole_matrix(carts) - (D + t[k] * S)).max() < 1e-12
    with pytest.raises(ValueError):
        ints.get_multipole_matrix(np.array([1, 1, 0]))


def test_integral_manager_exposes_dipoles():
    from slowquant_amd.integrals import IntegralManager

    sq = _h2()
    im = IntegralManager(sq)
    dx, dy, dz = im.electric_dipole
    assert dx.shape == (2, 2)
    assert np.allclose(np.diag(dz), sq.molecule.coordinates[:, 2])
