"""Synthetic inputs: determinism, conventions of the reference, physical invariants."""
import numpy as np

from dmftpack_amd import lattice as L


def test_prng_is_deterministic_and_uniform():
    a = L.uniform(7, 3, 10000)
    b = L.uniform(7, 3, 10000)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, L.uniform(7, 4, 10000))
    assert 0.0 <= a.min() and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.01
    z = L.normal(1, 1, 20000)
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03


def test_kmesh_matches_reference_convention():
    # TB.cpp:196-206 with a = pi: phase_j = -pi + pi/N + 2 pi j/N, k = x + y Nx + z Nx Ny
    th = L.kmesh_phases(4, 3, 2)
    assert th.shape == (24, 3)
    assert np.allclose(th[0], [-np.pi + np.pi / 4, -np.pi + np.pi / 3, -np.pi + np.pi / 2])
    assert np.allclose(th[1, 0] - th[0, 0], 2 * np.pi / 4)
    assert np.allclose(th[4, 1] - th[0, 1], 2 * np.pi / 3)
    assert np.allclose(th[12, 2] - th[0, 2], 2 * np.pi / 2)
    # symmetric mesh: sum of cosines vanishes
    assert abs(np.cos(L.kmesh_phases(16)[:, 0]).sum()) < 1e-10


def test_random_tb_is_hermitian_and_seeded():
    H = L.random_tb_hk(14, 4, seed=3)
    assert H.shape == (64, 14, 14)
    assert np.max(np.abs(H - np.conj(np.swapaxes(H, 1, 2)))) < 1e-14
    assert np.array_equal(H, L.random_tb_hk(14, 4, seed=3))
    # H(R) sum: the k average is the on-site block
    Rs, HR = L.random_tb_hr(14, seed=3)
    assert np.max(np.abs(H.mean(0) - HR[0])) < 1e-12
    eps = np.real(np.diag(HR[0]))
    assert eps.min() >= -3.0 and eps.max() <= 2.0


def test_pole_sigma_is_causal_with_exact_moments():
    sig, mom = L.pole_sigma(5, 64, 50.0, seed=2)
    # anti-Hermitian part negative semidefinite: Im Sigma <= 0
    for n in range(64):
        A = (sig[n] - sig[n].conj().T) / 2j
        assert np.linalg.eigvalsh(A).max() <= 1e-12
    # the moments are the exact 1/z expansion: check far out (beta small => w_n large)
    sig_hi, mom_hi = L.pole_sigma(5, 64, 0.05, seed=2)
    assert np.array_equal(mom_hi, mom)
    z = 1j * L.matsubara(64, 0.05)[-1]
    approx = mom[1] / z + mom[2] / z ** 2 + mom[3] / z ** 3
    assert np.max(np.abs(sig_hi[-1] - approx)) < 1e-9 * np.max(np.abs(sig_hi[-1]))


def test_t2g_dispersion():
    H, eps = L.t2g_hk(8)
    assert np.max(np.abs(H - np.einsum("ka,ab->kab", eps, np.eye(3)))) == 0
    assert abs(eps.mean()) < 1e-12
    assert np.max(np.abs(eps)) <= 4 * 0.25 + 0.03 + 1e-12
