"""Oracle checks for the callers either side of the lattice path: the
hybridisation function (Construct_hyb_delta.cpp:47-119), H(R) -> H(k)
(ConstructModelHamiltonian.cpp:58-77) and the Hk.HWR reader
(read_TB_parm.cpp:6-76).  Each is pinned by a closed form, not by retyping
the formula.

tests/golden/hubbard_Hk.HWR is the Hk.HWR input of DMFTpack's Hubbard-model
example (example/Hubb_model/primitive_Cell_ipt/Hk.HWR): one atom with two
orbitals and -0.25 eV hopping to the four in-plane neighbours.
"""
import os

import numpy as np

from dmftpack_amd import lattice as L
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
HWR = os.path.join(HERE, "golden", "hubbard_Hk.HWR")


def test_read_hwr_fixture():
    norbs, hops = L.read_hwr(HWR)
    assert norbs == [2]
    assert len(hops) == 8
    assert {h[3] for h in hops} == {0, 1} and all(h[3] == h[4] for h in hops)
    assert all(h[5] == -0.25 for h in hops)
    assert sorted((h[0], h[1], h[2]) for h in hops if h[3] == 0) == [(-1, 0, 0), (0, -1, 0), (0, 1, 0), (1, 0, 0)]


def test_hk_from_hr_hubbard_closed_form():
    """eps(k) = -2 * 0.25 (cos k_x a + cos k_y a) on both orbitals, no mixing."""
    _, hops = L.read_hwr(HWR)
    th = L.kmesh_phases(6, 6, 1)
    Hk = O.hk_from_hr(2, hops, th)
    eps = -0.5 * (np.cos(th[:, 0]) + np.cos(th[:, 1]))
    assert np.max(np.abs(Hk[:, 0, 0] - eps)) < 1e-15
    assert np.max(np.abs(Hk[:, 1, 1] - eps)) < 1e-15
    assert np.max(np.abs(Hk[:, 0, 1])) == 0 and np.max(np.abs(Hk[:, 1, 0])) == 0


def test_hk_from_hr_t2g_closed_form():
    th = L.kmesh_phases(5)
    Hk = O.hk_from_hr(3, L.t2g_hopping_list(), th)
    _, eps = L.t2g_hk(5)
    for a in range(3):
        assert np.max(np.abs(Hk[:, a, a] - eps[:, a])) < 1e-15
    off = Hk - np.einsum("kab,ab->kab", Hk, np.eye(3))
    assert np.max(np.abs(off)) == 0


def test_hk_from_hr_random_tb_matches_generator():
    """Full hopping list of the seeded d+p model -> the generator's H(k); Hermitian."""
    Rs, HR = L.random_tb_hr(6, 11)
    hops = L.hopping_list(Rs, HR)
    th = L.kmesh_phases(3)
    Hk = O.hk_from_hr(6, hops, th)
    ref = L.hk_from_hr(Rs, HR, 3)
    assert np.max(np.abs(Hk - ref)) < 1e-13
    assert np.max(np.abs(Hk - np.conj(np.swapaxes(Hk, 1, 2)))) < 1e-13


def hyb_case(nc=5, dim=3, nf=24, beta=20.0, mu=0.3, seed=3):
    """G_loc whose orbit sub-block is [(iw+mu) - H_imp - Sigma_imp - Delta0]^{-1}
    for a known causal Delta0 = sum_p V_p V_p^H / (iw - e_p)."""
    rng = np.random.default_rng(seed)
    orbit = rng.permutation(nc)[:dim]
    iw = 1j * np.pi * (2 * np.arange(nf) + 1) / beta
    X = rng.normal(size=(nc, nc)) + 1j * rng.normal(size=(nc, nc))
    h_imp = (X + X.conj().T) / 2
    sig, _ = L.pole_sigma(nc, nf, beta, seed)
    delta0 = np.zeros((nf, dim, dim), complex)
    for p in range(3):
        V = rng.normal(size=(dim, 2)) + 1j * rng.normal(size=(dim, 2))
        e = rng.uniform(-2, 2)
        delta0 += (V @ V.conj().T)[None] / (iw - e)[:, None, None]
    gloc = rng.normal(size=(nf, nc, nc)) + 1j * rng.normal(size=(nf, nc, nc))
    sub = np.ix_(orbit, orbit)
    for n in range(nf):
        A = (iw[n] + mu) * np.eye(dim) - h_imp[sub] - sig[n][sub] - delta0[n]
        gloc[n][sub] = np.linalg.inv(A)
    return gloc, sig, h_imp, orbit, beta, mu, delta0


def test_hyb_delta_round_trip():
    for nc, dim in ((5, 3), (5, 5), (10, 1), (10, 7)):
        gloc, sig, h_imp, orbit, beta, mu, delta0 = hyb_case(nc, dim)
        d = O.hyb_delta(gloc, sig, h_imp, orbit, beta, mu)
        assert np.max(np.abs(d - delta0)) <= 1e-12 * np.max(np.abs(delta0)) * 10, (nc, dim)


def test_hyb_delta_scalar_closed_form():
    """One orbital: Delta = iw + mu - h - Sigma - 1/G."""
    nf, beta, mu = 16, 10.0, -0.2
    iw = 1j * np.pi * (2 * np.arange(nf) + 1) / beta
    g = 1.0 / (iw + 0.7)
    s = 0.3 / (iw - 0.1)
    d = O.hyb_delta(g[:, None, None], s[:, None, None], np.array([[0.25]]), [0], beta, mu)
    ref = iw + mu - 0.25 - s - (iw + 0.7)
    assert np.max(np.abs(d[:, 0, 0] - ref)) < 1e-14


def t2g_real_sigma(ne, e0, de, seed=2):
    """Causal diagonal real-axis self-energy S_inf + A/(E - b + i Gamma)."""
    rng = np.random.default_rng(seed)
    E = e0 + de * np.arange(ne)
    sig = np.zeros((ne, 3, 3), complex)
    for a in range(3):
        s0, A, b, gam = rng.uniform(0, 1), rng.uniform(0.2, 1), rng.uniform(-1, 1), rng.uniform(0.2, 0.6)
        sig[:, a, a] = s0 + A / (E - b + 1j * gam)
    return sig


def test_greenftn_real_t2g_closed_form():
    """Orbital-diagonal H and Sigma: G^R_a(E) = mean_k 1/(E + mu + i eta - eps_a(k) - Sigma_a(E))."""
    Hk, eps = L.t2g_hk(6)
    ne, e0, de, eta, mu = 40, -2.0, 0.1, 0.05, 0.2
    sig = t2g_real_sigma(ne, e0, de)
    g, tr = O.greenftn_real(Hk, sig, mu, e0, de, eta)
    E = e0 + de * np.arange(ne)
    for a in range(3):
        ref = np.mean(1.0 / (E[:, None] + mu + 1j * eta - eps[None, :, a] - sig[:, a, a][:, None]), axis=1)
        assert np.max(np.abs(g[:, a, a] - ref)) < 1e-13
    assert np.max(np.abs(tr - np.trace(g, axis1=1, axis2=2))) < 1e-13


def test_greenftn_real_brute_force_and_sum_rule():
    """d+p model: numpy inverse on a few points; with Sigma = 0 the spectral
    weight -Im tr G / pi integrates to n (Lorentzian tails outside the window
    counted analytically)."""
    Hk = L.random_tb_hk(5, 2, 7)
    ne, e0, de, eta, mu = 3000, -15.0, 0.01, 0.05, 0.0
    sig = np.zeros((ne, 2, 2), complex)
    g, tr = O.greenftn_real(Hk, sig, mu, e0, de, eta, nthreads=4)
    for j in (0, 777, 2999):
        z = e0 + j * de + mu + 1j * eta
        G = np.mean([np.linalg.inv(z * np.eye(5) - H) for H in Hk], axis=0)
        assert np.max(np.abs(g[j] - G[:2, :2])) < 1e-13
        assert abs(tr[j] - np.trace(G)) < 1e-13
    ev = np.concatenate([np.linalg.eigvalsh(H) for H in Hk])
    lo, hi = e0 - de / 2, e0 + (ne - 0.5) * de
    # weight of each Lorentzian inside [lo, hi]
    inside = np.mean([np.sum((np.arctan((hi - x) / eta) - np.arctan((lo - x) / eta)) / np.pi)
                      for x in ev.reshape(len(Hk), -1)])
    weight = np.sum(-tr.imag / np.pi) * de
    assert abs(weight - inside) < 2e-3 * 5
    assert abs(inside - 5) < 0.02
