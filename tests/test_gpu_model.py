"""GPU hybridisation function and H(R) -> H(k) builder against the oracle.

Bar: 1e-12 normwise relative (hybridisation: one small inverse per frequency,
no pivoting on the GPU against the oracle's partial-pivot LU), 1e-13 absolute
for H(k) (a sum of a few exp terms per element).
"""
import os

import numpy as np
import pytest

from dmftpack_amd import lattice as L
from oracle import oracle as O
from test_cpu_model import hyb_case

HWR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hubbard_Hk.HWR")


@pytest.fixture(scope="module")
def gpu():
    from dmftpack_amd import gpu as G
    h = G.Handle()
    yield G, h
    h.close()


def rel(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.mark.gpu
@pytest.mark.parametrize("nc,dim", [(5, 1), (5, 3), (5, 5), (10, 7), (10, 10), (16, 13), (20, 17), (32, 32)])
def test_hyb_delta(gpu, nc, dim):
    G, h = gpu
    gloc, sig, h_imp, orbit, beta, mu, delta0 = hyb_case(nc, dim, nf=37)
    print(f"hyb_delta nc={nc} dim={dim}", flush=True)
    ref = O.hyb_delta(gloc, sig, h_imp, orbit, beta, mu)
    got = h.hyb_delta(gloc, sig, h_imp, orbit, beta, mu)
    assert rel(got, ref) <= 1e-12
    assert rel(got, delta0) <= 1e-11


@pytest.mark.gpu
def test_hyb_delta_from_lattice_gloc(gpu):
    """The chain the DMFT loop runs: G_loc from the lattice path -> Delta of the
    5 d orbitals; causal: -Im Delta(iw_n) is positive semidefinite."""
    G, h = gpu
    w = L.workload("dp14", seed=1, nk_side=4, n_freq=64)
    mom = O.asymto_moments(w["sigma"], w["beta"])
    lat = G.LatticeGF(w["Hk"], 5, 64, w["beta"], mom)
    out = lat.greenftn_w(w["sigma"], 0.3, method=G.SCHUR_F64)
    lat.close()
    h_imp = np.mean(w["Hk"], axis=0)[:5, :5]
    got = h.hyb_delta(out["gloc"], w["sigma"], h_imp, range(5), w["beta"], 0.3)
    ref = O.hyb_delta(out["gloc"], w["sigma"], h_imp, list(range(5)), w["beta"], 0.3)
    assert rel(got, ref) <= 1e-12
    for n in range(64):
        anti = (got[n] - got[n].conj().T) / 2j
        assert np.max(np.linalg.eigvalsh(anti)) <= 1e-9


@pytest.mark.gpu
def test_hyb_delta_errors(gpu):
    G, h = gpu
    gloc, sig, h_imp, orbit, beta, mu, _ = hyb_case(5, 3)
    with pytest.raises(G.capi.DmftGpuError):
        h.hyb_delta(gloc, sig, h_imp, [0, 1, 7], beta, mu)


@pytest.mark.gpu
def test_hk_from_hr_random(gpu):
    G, h = gpu
    Rs, HR = L.random_tb_hr(14, 5)
    hops = L.hopping_list(Rs, HR)
    th = L.kmesh_phases(6)
    ref = O.hk_from_hr(14, hops, th)
    dev = h.hk_from_hr(14, hops, th)
    got = G.from_dev_cm(dev)
    assert np.max(np.abs(got - ref)) <= 1e-13


@pytest.mark.gpu
def test_hk_from_hr_feeds_lattice(gpu):
    """H(k) built on the GPU from the Hk.HWR fixture goes straight into the
    lattice path; G_loc agrees with the oracle on the same H(k)."""
    G, h = gpu
    _, hops = L.read_hwr(HWR)
    th = L.kmesh_phases(8, 8, 1)
    Hk_dev = h.hk_from_hr(2, hops, th)
    Hk = O.hk_from_hr(2, hops, th)
    sig, _ = L.pole_sigma(2, 32, 40.0, 3)
    mom = O.asymto_moments(sig, 40.0)
    ref = O.greenftn_w(Hk, sig, mom, 0.1, 40.0)
    lat = G.LatticeGF(Hk_dev, 2, 32, 40.0, mom)
    out = lat.greenftn_w(sig, 0.1, method=G.PLAIN_F64)
    assert rel(out["gloc"], ref["gloc"]) <= 1e-12
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,method,tol", [("t2g", 0, 1e-12), ("dp14", 0, 1e-12), ("dp14", 1, 1e-12),
                                             ("dp14", 2, 1e-9), ("large32", 1, 1e-12), ("large32", 2, 1e-9)])
def test_greenftn_real(gpu, name, method, tol):
    """retarded_GreenFtn2 summed over k on a real-energy grid against the
    oracle; eta = 0.05 (the reference's infinitesimal), so |G| <= 20."""
    G, _ = gpu
    if name == "t2g":
        Hk, _ = L.t2g_hk(6)
        nc = 3
    elif name == "dp14":
        Hk, nc = L.random_tb_hk(14, 3, 4), 5
    else:
        Hk, nc = L.random_tb_hk(32, 2, 4), 10
    ne, e0, de, eta, mu = 150, -4.0, 0.05, 0.05, 0.1
    rng = np.random.default_rng(1)
    E = e0 + de * np.arange(ne)
    sig = np.zeros((ne, nc, nc), complex)
    for p in range(2):
        B = rng.normal(size=(nc, nc)) + 1j * rng.normal(size=(nc, nc))
        sig += (B @ B.conj().T / nc)[None] / (E - rng.uniform(-1, 1) + 0.3j)[:, None, None]
    print(f"greenftn_real {name} method {method}", flush=True)
    ref, tr_ref = O.greenftn_real(Hk, sig, mu, e0, de, eta, nthreads=8)
    mom = np.zeros((4, nc, nc), complex)
    lat = G.LatticeGF(Hk, nc, 16, 40.0, mom)
    got, tr = lat.greenftn_real(sig, mu, e0, de, eta, method=method)
    assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(tr - tr_ref)) <= tol * max(1.0, np.max(np.abs(tr_ref)))
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,nc", [(3, 3), (14, 5), (9, 2), (12, 6)])
@pytest.mark.parametrize("method,tol", [(1, 1e-12), (2, 1e-9)])
def test_greenftn_real_layouts(gpu, n, nc, method, tol):
    """The real axis through both Schur layouts for n_c <= 6 (one matrix per lane,
    the default, and one column per lane), k sums and per-k spectra, against the
    oracle."""
    G, _ = gpu
    Hk = L.t2g_hk(4)[0] if n == 3 else L.random_tb_hk(n, 3, 5)
    ne, e0, de, eta, mu = 130, -4.0, 0.06, 0.05, 0.2
    rng = np.random.default_rng(7)
    E = e0 + de * np.arange(ne)
    B = rng.normal(size=(nc, nc)) + 1j * rng.normal(size=(nc, nc))
    sig = (B @ B.conj().T / nc)[None] / (E - 0.3 + 0.5j)[:, None, None]
    ref, tr_ref = O.greenftn_real(Hk, sig, mu, e0, de, eta, nthreads=8)
    print(f"greenftn_real layouts n={n} nc={nc} method {method}", flush=True)
    lat = G.LatticeGF(Hk, nc, 16, 40.0, np.zeros((4, nc, nc), complex))
    outs = []
    for layout in (0, 1):
        lat.set_option(G.capi.OPT_SCHUR_LAYOUT, layout)
        got, tr = lat.greenftn_real(sig, mu, e0, de, eta, method=method)
        assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref))), layout
        assert np.max(np.abs(tr - tr_ref)) <= tol * max(1.0, np.max(np.abs(tr_ref))), layout
        gk, trk = lat.greenftn_real_k(sig, mu, e0, de, eta, method=method)
        assert np.max(np.abs(gk.mean(axis=0) - got)) <= 1e-13 * max(1.0, np.max(np.abs(got)))
        outs.append(gk)
    ref0, _ = O.greenftn_real(Hk[:1], sig, mu, e0, de, eta)
    for gk in outs:
        assert np.max(np.abs(gk[0] - ref0)) <= tol * max(1.0, np.max(np.abs(ref0)))
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,method,tol", [("t2g", 0, 1e-12), ("dp14", 1, 1e-12), ("dp14", 2, 1e-9),
                                             ("large32", 1, 1e-12)])
def test_greenftn_real_per_k(gpu, name, method, tol):
    """Band spectra on a k path (qs_spectra.cpp:197-221): the per-k retarded G
    against the oracle run on one k at a time, and its k average against the
    summed entry point."""
    G, _ = gpu
    if name == "t2g":
        Hk, _ = L.t2g_hk(4)
        nc = 3
    elif name == "dp14":
        Hk, nc = L.random_tb_hk(14, 2, 4), 5
    else:
        Hk, nc = L.random_tb_hk(32, 2, 4), 10
    Hk = Hk[:7]  # a short k path
    ne, e0, de, eta, mu = 90, -3.0, 0.07, 0.05, 0.1
    rng = np.random.default_rng(3)
    E = e0 + de * np.arange(ne)
    B = rng.normal(size=(nc, nc)) + 1j * rng.normal(size=(nc, nc))
    sig = (B @ B.conj().T / nc)[None] / (E - 0.2 + 0.4j)[:, None, None]
    print(f"greenftn_real_k {name} method {method}", flush=True)
    lat = G.LatticeGF(Hk, nc, 16, 40.0, np.zeros((4, nc, nc), complex))
    gk, trk = lat.greenftn_real_k(sig, mu, e0, de, eta, method=method)
    assert gk.shape == (7, ne, nc, nc) and trk.shape == (7, ne)
    for k in range(7):
        ref, tr_ref = O.greenftn_real(Hk[k:k + 1], sig, mu, e0, de, eta)
        assert np.max(np.abs(gk[k] - ref)) <= tol * max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(trk[k] - tr_ref)) <= tol * max(1.0, np.max(np.abs(tr_ref)))
    g_avg, tr_avg = lat.greenftn_real(sig, mu, e0, de, eta, method=method)
    assert np.max(np.abs(gk.mean(axis=0) - g_avg)) <= 1e-13 * max(1.0, np.max(np.abs(g_avg)))
    assert np.max(np.abs(trk.mean(axis=0) - tr_avg)) <= 1e-13 * max(1.0, np.max(np.abs(tr_avg)))
    # spectral function of the band plot: A(k, E) = -Im tr G^R / pi >= 0
    assert np.min(-trk.imag) > -1e-12
    lat.close()
