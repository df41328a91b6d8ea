"""Pins the FP64 oracle against the reference's own known-answer tests.

Each test restates one reference test (file:line in the muspinsim tree) as a
check of the oracle; the expected values are the reference's literals or
closed forms, never the oracle's own formula retyped.
"""

import numpy as np
import pytest
import scipy.constants as cnst

from oracle import dynamics, orient, physics, spin


def test_pauli_and_spin1():
    """tests/test_spinop.py:41-62."""
    assert np.all(spin.s_x(0.5) == [[0, 0.5], [0.5, 0]])
    assert np.all(spin.s_y(0.5) == [[0, -0.5j], [0.5j, 0]])
    assert np.all(spin.s_z(0.5) == [[0.5, 0], [0, -0.5]])
    assert np.all(spin.s_z(1) == np.diag([1, 0, -1]))
    assert np.allclose(spin.s_x(1), np.array([[0, 1, 0.0], [1, 0, 1.0], [0, 1, 0.0]]) / 2 ** 0.5)


def test_commutators_and_kron():
    """tests/test_spinop.py:64-103."""
    sx, sy, sz = spin.s_x(0.5), spin.s_y(0.5), spin.s_z(0.5)
    assert np.all(sx @ sy - sy @ sx == 1.0j * sz)
    assert np.all(sy @ sz - sz @ sy == 1.0j * sx)
    assert np.all(sz @ sx - sx @ sz == 1.0j * sy)
    assert np.all(4 * np.kron(sx, sz) == [[0, 0, 1, 0], [0, 0, 0, -1], [1, 0, 0, 0], [0, -1, 0, 0]])
    # Hilbert-Schmidt with a state along (1,1,0)/sqrt2
    rho = spin.pure_state(0.5, np.array([1, 1, 0]) / 2 ** 0.5)
    assert np.isclose(2 * np.real(np.trace(rho.conj().T @ sx)), 0.5 ** 0.5)


@pytest.mark.parametrize("I", [0.5, 1.0, 1.5, 3.5])
def test_spin_algebra_general(I):
    """[Sx, Sy] = i Sz and S^2 = I(I+1) for any spin (textbook identities)."""
    sx, sy, sz = spin.s_x(I), spin.s_y(I), spin.s_z(I)
    assert np.allclose(sx @ sy - sy @ sx, 1j * sz)
    assert np.allclose(sx @ sx + sy @ sy + sz @ sz, I * (I + 1) * np.eye(len(sz)))


def test_multi_spin_product():
    """tests/test_spinop.py:149-163: 16 SxSx SySy = diag(-1, 1, 1, -1)."""
    SxSx = np.kron(spin.s_x(0.5), spin.s_x(0.5))
    SySy = np.kron(spin.s_y(0.5), spin.s_y(0.5))
    assert np.allclose(16 * SxSx @ SySy, np.diag([-1, 1, 1, -1]))


def test_density_from_vectors():
    """tests/test_spinop.py:165-203."""
    assert np.all(spin.pure_state(0.5, [1, 0, 0]) == np.ones((2, 2)) * 0.5)
    r = spin.from_vectors([0.5], [[0, 1, 0]], [0.5])
    assert np.allclose(r, np.array([[0.5, -0.25j], [0.25j, 0.5]]))


def test_superoperators():
    """tests/test_spinop.py:225-245 on row-major vec."""
    sx = spin.s_x(0.5)
    rho = spin.pure_state(0.5, [0, 0, 1])
    v = rho.reshape(-1)
    assert np.allclose((spin.left_mult(sx) @ v).reshape(2, 2), sx @ rho)
    assert np.allclose((spin.commutator(sx) @ v).reshape(2, 2), sx @ rho - rho @ sx)
    assert np.allclose((spin.anticommutator(sx) @ v).reshape(2, 2), sx @ rho + rho @ sx)
    assert np.allclose((spin.bracket(sx) @ v).reshape(2, 2), sx @ rho @ sx.conj().T)


def test_hamiltonian_addterms():
    """tests/test_spinsys.py:136-180: exact 4x4 H for (mu, e) with linear and
    bilinear terms."""
    Is = [0.5, 0.5]
    terms = [("linear", 0, 0, [1, 0, 1]), ("bilinear", 0, 1, np.eye(3))]
    H = physics.hamiltonian(Is, terms)
    expect = np.array([[0.75, 0, 0.5, 0.0], [0, 0.25, 0.5, 0.5],
                       [0.5, 0.5, -0.75, 0], [0.0, 0.5, 0, -0.25]])
    assert np.allclose(H, expect)


def test_bilinear_rotation():
    """tests/test_spinsys.py:47-76: R M R^T with M = diag(1,1,2)."""
    Is = [0.5, 0.5]
    M = np.diag([1, 1, 2.0])
    R = np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]])
    H = physics.hamiltonian(Is, [("bilinear", 0, 1, R @ M @ R.T)])
    ops = {a: np.kron(spin.single(0.5, a), spin.single(0.5, a)) for a in "xyz"}
    assert np.allclose(H, ops["x"] + 2 * ops["y"] + ops["z"])


def test_zeeman_hyperfine_diagonal():
    """tests/test_spinsys.py:364-413 (e, mu with Zeeman at 1 T + 100 MHz iso)."""
    gmu, ge = physics.MU_GAMMA, physics.ELEC_GAMMA
    Is = [0.5, 0.5]
    H = physics.zeeman(Is, [ge, gmu], [0, 0, 1.0]) + physics.hamiltonian(
        Is, [("bilinear", 1, 0, np.eye(3) * 100)])
    assert np.allclose(np.diag(H), [0.5 * (ge + gmu) + 25, 0.5 * (ge - gmu) - 25,
                                    0.5 * (-ge + gmu) - 25, 0.5 * (-ge - gmu) + 25])


def test_constants():
    """tests/test_constants.py:7-12."""
    assert np.isclose(physics.ELEC_GAMMA, -28024.9514242)
    # assertAlmostEqual (7 places) as in the reference; EFG_2_MHZ moves in the
    # 10th digit between CODATA releases shipped with scipy.
    assert round(physics.MU_GAMMA - 135.53880943285955, 7) == 0
    assert physics.MU_TAU == 2.19703
    assert round(physics.EFG_2_MHZ - 0.23496477815245767, 7) == 0


def test_muon_operator_and_sigma():
    """tests/test_spinsys.py:421-448."""
    op = physics.muon_operator([0.5, 0.5], 0, [1, 1, 1])
    assert np.array_equal(op, [[0.5, 0, 0.5 - 0.5j, 0], [0, 0.5, 0, 0.5 - 0.5j],
                               [0.5 + 0.5j, 0, -0.5, 0], [0, 0.5 + 0.5j, 0, -0.5]])
    assert np.array_equal(physics.sigma_mu([1, 1, 1]), [[1, 1 - 1j], [1 + 1j, -1]])


def test_rho0_infinite_T_and_boltzmann():
    """tests/test_experiment.py:74-115."""
    ge, gmu = physics.ELEC_GAMMA, physics.MU_GAMMA
    r = physics.rho0([0.5, 0.5], [ge, gmu], 1, [0, 0, 0], [1, 0, 0], np.inf)
    assert np.allclose(r, [[0.25, 0.25, 0, 0], [0.25, 0.25, 0, 0],
                           [0, 0, 0.25, 0.25], [0, 0, 0.25, 0.25]])
    T = 100
    B = [0, 0, 2.0e-6 * cnst.k * T / (ge * cnst.h)]
    r = physics.rho0([0.5, 0.5], [ge, gmu], 1, B, [0, 0, 1.0], T)
    Z = np.exp([-1, 1])
    Z /= np.sum(Z)
    assert np.allclose(np.diag(r), [Z[0], 0, Z[1], 0])


def test_eigh_known_answer():
    """tests/test_hamiltonian.py:20-32 (Sx -> E = [-1/2, 1/2])."""
    E, U = dynamics.eigh(spin.s_x(0.5))
    assert np.all(E == [-0.5, 0.5])
    ref = np.array([[1.0, 1.0], [-1.0, 1.0]]) / 2 ** 0.5
    assert np.allclose(abs(U @ ref), np.eye(2))


def test_evolve_precession():
    """tests/test_hamiltonian.py:34-45: 0.5 cos(2 pi t)."""
    H = physics.hamiltonian([0.5], [("linear", 0, 0, [1, 0, 0])])
    rho0 = spin.pure_state(0.5, [0, 0, 1])
    t = np.linspace(0, 1, 100)
    ev = dynamics.evolve(H, rho0, t, [spin.s_z(0.5)])
    assert np.allclose(ev[:, 0], 0.5 * np.cos(2 * np.pi * t))


def test_evolve_density_four_times():
    """tests/test_hamiltonian.py:47-103."""
    H = physics.hamiltonian([0.5], [("linear", 0, 0, [1, 0, 0])])
    rho0 = spin.pure_state(0.5, [0, 0, 1])
    ev = dynamics.evolve_density(H, rho0, np.linspace(0, 1, 4))
    assert np.allclose(ev[0], [[1, 0], [0, 0]])
    assert np.allclose(ev[1], [[0.25, 0.433012702j], [-0.433012702j, 0.75]])
    assert np.allclose(ev[2], [[0.25, -0.433012702j], [0.433012702j, 0.75]])
    assert np.allclose(ev[3], [[1, 0], [0, 0]])


def test_fast_evolve_precession():
    """tests/test_hamiltonian.py:105-116."""
    H = physics.hamiltonian([0.5, 0.5], [("linear", 0, 0, [1, 0, 0])])
    t = np.linspace(0, 1, 100)
    ev = dynamics.fast_evolve(H, physics.sigma_mu([0, 0, 1]), t, 2)
    assert np.allclose(ev, 0.5 * np.cos(2 * np.pi * t))


def test_integrate_closed_form():
    """tests/test_hamiltonian.py:118-128: 0.5 / (1 + 4 pi^2)."""
    H = physics.hamiltonian([0.5], [("linear", 0, 0, [1, 0, 0])])
    rho0 = spin.pure_state(0.5, [0, 0, 1])
    avg = dynamics.integrate_decaying(H, rho0, 1.0, [spin.s_z(0.5)])
    assert np.isclose(avg[0], 0.5 / (1.0 + 4 * np.pi ** 2))


def test_lindbladian_matrices():
    """tests/test_lindbladian.py:10-38."""
    sx, sz = spin.s_x(0.5), spin.s_z(0.5)
    L = physics.lindbladian(sz, [])
    assert np.all(L == 1.0j * np.diag([0, -1, 1, 0]))
    L = physics.lindbladian(sz, [(sx, 0.1)])
    assert np.allclose(L, [[-0.025, 0, 0, 0.025], [0, -0.025 - 1.0j, 0.025, 0],
                           [0, 0.025, -0.025 + 1.0j, 0], [0.025, 0, 0, -0.025]])


def test_lindblad_evolve_closed_forms():
    """tests/test_lindbladian.py:40-80 (damped precession, S+/S- rates)."""
    sx, sz = spin.s_x(0.5), spin.s_z(0.5)
    sp, sm = spin.s_plus(0.5), spin.s_minus(0.5)
    rho0 = spin.pure_state(0.5, [1, 0, 0])
    t = np.linspace(0, 1, 100)
    ev = dynamics.lindblad_evolve(physics.lindbladian(sz, []), rho0, t, [sx])
    assert np.allclose(ev[:, 0], 0.5 * np.cos(2 * np.pi * t))
    for g in [1.0, 2.0, 5.0, 10.0]:
        ev = dynamics.lindblad_evolve(physics.lindbladian(sz, [(sx, g)]), rho0, t, [sx])
        ap = -0.5 * np.pi * g + ((0.5 * np.pi * g) ** 2 - 4 * np.pi ** 2 + 0j) ** 0.5
        am = -0.5 * np.pi * g - ((0.5 * np.pi * g) ** 2 - 4 * np.pi ** 2 + 0j) ** 0.5
        A = ap * am / (am - ap)
        solx = np.real(0.5 * A * (np.exp(ap * t) / ap - np.exp(am * t) / am))
        assert np.allclose(ev[:, 0], solx)
        ev = dynamics.lindblad_evolve(physics.lindbladian(sz, [(sp, 1.5 * g), (sm, 0.5 * g)]),
                                      rho0, t, [sx, sz])
        assert np.allclose(ev[:, 0], np.real(0.5 * np.cos(2 * np.pi * t) * np.exp(-2 * np.pi * g * t)))
        assert np.allclose(ev[:, 1], 0.25 * (1 - np.exp(-4 * np.pi * g * t)))


def test_lindblad_integral_closed_forms():
    """tests/test_lindbladian.py:130-159."""
    sx, sz = spin.s_x(0.5), spin.s_z(0.5)
    rho0 = spin.pure_state(0.5, [1, 0, 0])
    tau = 2.0
    avg = dynamics.lindblad_integrate_decaying(physics.lindbladian(sz, []), rho0, tau, [sx])
    assert np.isclose(avg[0], 0.5 * tau / (1 + 4 * np.pi ** 2 * tau ** 2))
    for g in [1.0, 2.0, 5.0, 10.0]:
        avg = dynamics.lindblad_integrate_decaying(physics.lindbladian(sz, [(sx, g)]), rho0, tau, [sx])
        ap = -0.5 * np.pi * g + ((0.5 * np.pi * g) ** 2 - 4 * np.pi ** 2 + 0j) ** 0.5
        am = -0.5 * np.pi * g - ((0.5 * np.pi * g) ** 2 - 4 * np.pi ** 2 + 0j) ** 0.5
        A = ap * am / (am - ap)
        sol = np.real(0.5 * A * tau * (1 / ((1 - ap * tau) * ap) - 1 / ((1 - am * tau) * am)))
        assert np.isclose(avg[0], sol)


def test_quaternion_conventions():
    """tests/test_utils.py:43-65 and tests/test_config.py:214-286."""
    z1 = orient.rotate(orient.quat_from_polar(np.pi / 4.0, 0), [0, 0, 1])
    assert np.allclose(z1, [2 ** -0.5, 0, 2 ** -0.5])
    th, ph = 0.6 * np.pi, 0.4 * np.pi
    q2 = orient.quat_from_polar(th, ph)
    assert np.allclose(orient.rotate(q2, [0, 0, 1]),
                       [np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    assert np.allclose(orient.rotate(orient.qconj(q2), [0, 0, 1]),
                       [-np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    # stored conjugate of (pi/2, 0, 0) zyz  (tests/test_config.py:226-231)
    qs, w = orient.orientation_entry([0.5 * np.pi, 0, 0, 2.0])
    assert np.allclose(qs, [2 ** -0.5, 0, 0, -(2 ** -0.5)]) and w == 2.0
    # rotation matrix of q agrees with q v q*
    q = orient.from_euler(0.3, 1.1, -0.7)
    v = np.array([0.2, -1.3, 0.5])
    assert np.allclose(orient.rotation_matrix(q) @ v, orient.rotate(q, v))


def test_zcw_and_eulrange():
    """tests/test_utils.py:67-94."""
    o = orient.zcw_angles(1000)
    assert len(o) >= 1000
    assert abs(np.average(3.0 * np.cos(o[:, 0]) ** 2 - 1.0)) < 5e-6
    ow = orient.eulrange(10)
    A = np.array([[1, 2, 3], [2, 4, 4], [3, 4, -3]])
    Asum = np.zeros((3, 3))
    for a, b, c, w in ow:
        R = orient.rotation_matrix(orient.from_euler(a, b, c))
        Asum += (R @ A @ R.T) * w
    Asum /= np.sum(ow[:, -1])
    assert np.sum(np.abs(Asum - np.eye(3) * np.trace(A) / 3)) < 1.0


def test_get_xy():
    """tests/test_utils.py:96-109."""
    x, y = physics.get_xy(np.array([0, 0, 1]))
    assert np.allclose(x, [1, 0, 0])
    z = np.array([1, 1, 2.0])
    x, y = physics.get_xy(z)
    assert np.allclose(np.cross(x, y), z / np.linalg.norm(z))
    with pytest.raises(ValueError):
        physics.get_xy([0, 0, 0])


def _dipolar_modules():
    from muspinsim_amd import constants as product_constants
    return [physics, product_constants]


@pytest.mark.parametrize("mod", _dipolar_modules(), ids=["oracle", "package"])
def test_dipolar_tensor_textbook(mod):
    """Dipolar tensor (spinsys.py:400-409) against the textbook form
    H_D = b [S_i.S_j - 3 (S_i.r)(S_j.r)], b = (mu0 / 4 pi) hbar g_i g_j / r^3 with
    the gyromagnetic ratios in rad s^-1 T^-1 and b converted from rad/s to MHz.
    For two protons 1 A apart b is the familiar 120.1 kHz.  The tensor is
    b (1 - 3 r r^T): traceless, eigenvalue -2b along r and +b twice across it."""
    gH = 42.577478518                       # MHz/T, pinned by tests/test_constants.py
    grad = 2 * np.pi * gH * 1e6             # rad s^-1 T^-1

    def b_mhz(g1, g2, r_ang):
        b = cnst.mu_0 / (4 * np.pi) * cnst.hbar * (2 * np.pi * g1 * 1e6) * (2 * np.pi * g2 * 1e6)
        return b / (r_ang * 1e-10) ** 3 / (2 * np.pi) / 1e6

    b1 = cnst.mu_0 / (4 * np.pi) * cnst.hbar * grad ** 2 / (1e-10) ** 3 / (2 * np.pi) / 1e6
    assert abs(b1 - 0.1201) < 5e-5
    rng = np.random.default_rng(17)
    pairs = [(gH, gH), (physics.MU_GAMMA, gH), (physics.ELEC_GAMMA, physics.MU_GAMMA)]
    for g1, g2 in pairs:
        for _ in range(3):
            r = rng.normal(size=3)
            r *= rng.uniform(0.8, 3.0) / np.linalg.norm(r)
            rn = np.linalg.norm(r)
            rh = r / rn
            b = b_mhz(g1, g2, rn)
            D = mod.dipolar_tensor(g1, g2, r)
            assert np.allclose(D, b * (np.eye(3) - 3 * np.outer(rh, rh)), rtol=1e-12, atol=0)
            assert abs(np.trace(D)) <= 1e-12 * abs(b)
            assert np.allclose(D @ rh, -2 * b * rh, rtol=1e-12, atol=1e-15 * abs(b))
            ev = np.sort(np.linalg.eigvalsh(D))
            assert np.allclose(ev, np.sort([-2 * b, b, b]), rtol=1e-12, atol=1e-15 * abs(b))
    # the operator: two protons on the z axis, H = b (S1.S2 - 3 S1z S2z)
    D = mod.dipolar_tensor(gH, gH, [0, 0, 1.5])
    H = physics.hamiltonian([0.5, 0.5], [("bilinear", 0, 1, D)])
    s = [spin.single(0.5, a) for a in "xyz"]
    SS = sum(np.kron(x, x) for x in s)
    b = b_mhz(gH, gH, 1.5)
    assert np.allclose(H, b * (SS - 3 * np.kron(s[2], s[2])), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("mod", _dipolar_modules(), ids=["oracle", "package"])
def test_quadrupolar_spin1_closed_form(mod):
    """Quadrupolar tensor (spinsys.py:433-438) through the spin-1 closed form.
    EFG_2_MHZ is rebuilt from CODATA as e (E_h / (e a0^2)) (1 mb) / h in MHz.
    For an EFG with principal values |Vzz| >= |Vyy| >= |Vxx| (atomic units) in
    a random frame and eta = (Vxx - Vyy) / Vzz, H_Q = sum_ab Q_ab S_a S_b has
    eigenvalues (Cq/4)(1 - eta), (Cq/4)(1 + eta) and -Cq/2 with
    Cq = EFG_2_MHZ Q Vzz; the spin-3/2 levels are +-(Cq/4) sqrt(1 + eta^2/3),
    each twice (Kramers pairs)."""
    Eh = cnst.physical_constants["Hartree energy"][0]
    a0 = cnst.physical_constants["Bohr radius"][0]
    efg_au = Eh / (cnst.e * a0 ** 2)                     # V / m^2
    k = cnst.e * efg_au * 1e-31 / cnst.h / 1e6           # MHz per (a.u. x mb)
    assert np.isclose(mod.EFG_2_MHZ, k, rtol=1e-8, atol=0)
    rng = np.random.default_rng(23)
    for _ in range(4):
        Vzz = rng.uniform(0.1, 2.0) * rng.choice([-1, 1])
        eta = rng.uniform(0, 1)
        Vxx = -Vzz * (1 - eta) / 2
        Vyy = -Vzz * (1 + eta) / 2
        Rq, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        V = Rq @ np.diag([Vxx, Vyy, Vzz]) @ Rq.T
        Q = rng.uniform(1.0, 50.0)
        Cq = k * Q * Vzz
        H = physics.hamiltonian([1.0], [("bilinear", 0, 0, mod.quadrupolar_tensor(Q, 1.0, V))])
        ev = np.sort(np.linalg.eigvalsh(H))
        ref = np.sort([Cq / 4 * (1 - eta), Cq / 4 * (1 + eta), -Cq / 2])
        assert np.allclose(ev, ref, rtol=1e-10, atol=1e-12 * abs(Cq))
        H = physics.hamiltonian([1.5], [("bilinear", 0, 0, mod.quadrupolar_tensor(Q, 1.5, V))])
        ev = np.sort(np.linalg.eigvalsh(H))
        e0 = abs(Cq) / 4 * np.sqrt(1 + eta ** 2 / 3)
        assert np.allclose(ev, [-e0, -e0, e0, e0], rtol=1e-10, atol=1e-12 * abs(Cq))
    with pytest.raises(ValueError):
        mod.quadrupolar_tensor(1.0, 0.5, np.eye(3))
