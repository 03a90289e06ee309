"""The lattice path at the full sizes of the benchmark configurations.

- t2g (3 orbitals, Nk=16^3, N_freq=1024, beta=40): GPU against the oracle
  outright (the oracle finishes in seconds at this size), plus the scalar
  closed form G_loc,a = <1/(iw + mu - eps_a(k) - Sigma_a(iw))>_k.
- dp14 (14 orbitals, n_c=5, Nk=32^3, N_freq=2048, beta=50): the oracle is too
  slow for the whole mesh, so the full-size run is checked through properties
  that do not depend on size: the three methods agree (plain FP64 = Schur FP64
  to 1e-12, FP32+refine to 1e-9), NumMatrix is Hermitian with eigenvalues in
  [0, 1] (GreenFtn.cpp:358-365), N equals the trace of NumMatrix plus the d
  block (checked through N(mu) of the density path), and on a k subset the
  GPU matches the oracle.
- dp14_grad (14 orbitals at Nk=16^3, N_freq=1024, beta=40): the adjoint at full
  size against central finite differences of the GPU loss (h=1e-6) in a few
  Sigma entries and in mu, relative 1e-6.
- large32 (32 orbitals, two 5x5 clusters): full frequency count on a reduced
  mesh against the oracle; Schur FP64 against FP32+refine at full mesh.
"""
import numpy as np
import pytest

from dmftpack_amd import lattice as L
from oracle import oracle as O


def rel(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="module")
def gpu():
    from dmftpack_amd import gpu as G
    return G


@pytest.mark.gpu
def test_t2g_full_size(gpu):
    w = L.workload("t2g", seed=1)
    Hk, sig, beta = w["Hk"], w["sigma"], w["beta"]
    mom = O.asymto_moments(sig, beta)
    print(f"t2g full: nk={Hk.shape[0]} nf={sig.shape[0]}", flush=True)
    ref = O.greenftn_w(Hk, sig, mom, 0.1, beta, nthreads=16)
    lat = gpu.LatticeGF(Hk, 3, sig.shape[0], beta, mom)
    for method, tol in ((gpu.PLAIN_F64, 1e-12), (gpu.SCHUR_F64, 1e-12), (gpu.SCHUR_F32R, 1e-9)):
        out = lat.greenftn_w(sig, 0.1, method=method)
        assert rel(out["gloc"], ref["gloc"]) <= tol, method
        assert rel(out["num_matrix"], ref["num_matrix"]) <= tol, method
    # closed form of the orbital-diagonal model
    _, eps = L.t2g_hk(16)
    iw = 1j * np.pi * (2 * np.arange(sig.shape[0]) + 1) / beta
    out = lat.greenftn_w(sig, 0.1, method=gpu.PLAIN_F64)
    for a in range(3):
        cf = np.mean(1.0 / (iw[:, None] + 0.1 - eps[None, :, a] - sig[:, a, a][:, None]), axis=1)
        assert np.max(np.abs(out["gloc"][:, a, a] - cf)) < 1e-13
    # mu search to the t2g filling, Newton on the analytic dN/dmu
    mu, nev = lat.adjust_mu(sig, 0.0, w["n_target"], method=gpu.SCHUR_F64, tol=1e-10)
    N, _ = lat.density(sig, mu, method=gpu.PLAIN_F64)
    assert abs(N - w["n_target"]) < 1e-9
    lat.close()


@pytest.mark.gpu
def test_dp14_full_size_properties(gpu):
    w = L.workload("dp14", seed=1)
    Hk, sig, beta, mom = w["Hk"], w["sigma"], w["beta"], w["exact_moments"]
    nf = sig.shape[0]
    print(f"dp14 full: nk={Hk.shape[0]} nf={nf}", flush=True)
    lat = gpu.LatticeGF(Hk, 5, nf, beta, mom)
    p = lat.greenftn_w(sig, 0.25, method=gpu.PLAIN_F64)
    s = lat.greenftn_w(sig, 0.25, method=gpu.SCHUR_F64)
    r = lat.greenftn_w(sig, 0.25, method=gpu.SCHUR_F32R)
    assert rel(s["gloc"], p["gloc"]) <= 1e-12 and rel(s["num_matrix"], p["num_matrix"]) <= 1e-12
    assert np.max(np.abs(r["gloc"] - p["gloc"])) <= 1e-9 and rel(r["num_matrix"], p["num_matrix"]) <= 1e-9
    assert abs(s["n_total"] - p["n_total"]) <= 1e-10 and abs(r["n_total"] - p["n_total"]) <= 1e-9
    nm = p["num_matrix"]
    assert np.max(np.abs(nm - nm.conj().T)) < 1e-12
    ev = np.linalg.eigvalsh(0.5 * (nm + nm.conj().T))
    assert ev.min() > -1e-9 and ev.max() < 1 + 1e-9
    # N of the forward pass equals N(mu) of the density path
    N, dN = lat.density(sig, 0.25, method=gpu.SCHUR_F64)
    assert abs(N - p["n_total"]) <= 1e-10 and dN > 0
    # G_loc(iw) of a causal lattice: -Im G_loc is positive definite
    for n in (0, nf // 2, nf - 1):
        anti = (p["gloc"][n] - p["gloc"][n].conj().T) / 2j
        assert np.max(np.linalg.eigvalsh(anti)) < 0
    lat.close()
    # the oracle on a k subset of the same lattice, all frequencies
    sub = Hk[::256]
    ref = O.greenftn_w(sub, sig, mom, 0.25, beta, nthreads=16)
    lat = gpu.LatticeGF(sub, 5, nf, beta, mom)
    for method, tol in ((gpu.PLAIN_F64, 1e-12), (gpu.SCHUR_F64, 1e-12), (gpu.SCHUR_F32R, 1e-9)):
        out = lat.greenftn_w(sig, 0.25, method=method)
        assert rel(out["gloc"], ref["gloc"]) <= tol, method
        assert rel(out["num_matrix"], ref["num_matrix"]) <= tol, method
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("method", [1, 2])
def test_dp14_grad_finite_differences(gpu, method):
    """dL/dSigma and dL/dmu of L = sum_w ||G_loc - G_tgt||_F^2 + (N - N0)^2 at the
    differentiation configuration's full size, against central differences."""
    w = L.workload("dp14_grad", seed=1)
    Hk, sig, beta, mom = w["Hk"], w["sigma"], w["beta"], w["exact_moments"]
    nf, mu, N0 = sig.shape[0], 0.25, w["n_target"]
    print(f"dp14_grad full: nk={Hk.shape[0]} nf={nf} method={method}", flush=True)
    lat = gpu.LatticeGF(Hk, 5, nf, beta, mom)
    tgt = lat.greenftn_w(L.perturbed(sig, seed=7), mu, method=gpu.SCHUR_F64)["gloc"]

    def loss(s, m):
        o = lat.greenftn_w(s, m, method=gpu.SCHUR_F64)
        return float(np.sum(np.abs(o["gloc"] - tgt) ** 2) + (o["n_total"] - N0) ** 2), o

    L0, o = loss(sig, mu)
    gbar = 2.0 * (o["gloc"] - tgt)
    sb, mb = lat.vjp(sig, mu, gbar, 2.0 * (o["n_total"] - N0), method=method)
    h = 1e-6
    rng = np.random.default_rng(3)
    for _ in range(4):
        n, a, b = int(rng.integers(nf)), int(rng.integers(5)), int(rng.integers(5))
        for part, unit in (("re", 1.0), ("im", 1j)):
            sp, sm = sig.copy(), sig.copy()
            sp[n, a, b] += h * unit
            sm[n, a, b] -= h * unit
            fd = (loss(sp, mu)[0] - loss(sm, mu)[0]) / (2 * h)
            an = sb[n, a, b].real if part == "re" else sb[n, a, b].imag
            assert abs(an - fd) <= 1e-6 * max(abs(fd), np.max(np.abs(sb))), (n, a, b, part, an, fd)
    fd_mu = (loss(sig, mu + h)[0] - loss(sig, mu - h)[0]) / (2 * h)
    assert abs(mb - fd_mu) <= 1e-6 * max(1.0, abs(fd_mu))
    lat.close()


@pytest.mark.gpu
def test_large32_full_frequencies(gpu):
    """32 orbitals, two 5x5 clusters, beta=100, all 4096 frequencies: oracle on
    a 4^3 mesh; Schur FP64 vs FP32+refine on the full 40^3 mesh."""
    w = L.workload("large32", seed=1, nk_side=4)
    Hk, sig, beta = w["Hk"], w["sigma"], w["beta"]
    mom = w["exact_moments"]
    nf = sig.shape[0]
    print(f"large32 small mesh: nk={Hk.shape[0]} nf={nf}", flush=True)
    ref = O.greenftn_w(Hk, sig, mom, 0.0, beta, n_cluster=2, nthreads=16)
    lat = gpu.LatticeGF(Hk, 10, nf, beta, mom, n_cluster=2)
    for method, tol in ((gpu.SCHUR_F64, 1e-12), (gpu.SCHUR_F32R, 1e-9)):
        out = lat.greenftn_w(sig, 0.0, method=method)
        assert rel(out["gloc"], ref["gloc"]) <= tol, method
        assert rel(out["num_matrix"], ref["num_matrix"]) <= tol, method
    lat.close()
    w = L.workload("large32", seed=1)
    print(f"large32 full mesh: nk={w['Hk'].shape[0]}", flush=True)
    lat = gpu.LatticeGF(w["Hk"], 10, nf, beta, mom, n_cluster=2)
    del w
    s = lat.greenftn_w(sig, 0.0, method=gpu.SCHUR_F64)
    r = lat.greenftn_w(sig, 0.0, method=gpu.SCHUR_F32R)
    assert np.max(np.abs(r["gloc"] - s["gloc"])) <= 1e-9
    assert rel(r["num_matrix"], s["num_matrix"]) <= 1e-9
    # inter-cluster blocks of G_loc are zero (GreenFtn.cpp:371-383)
    assert np.max(np.abs(s["gloc"][:, :5, 5:])) == 0 and np.max(np.abs(s["gloc"][:, 5:, :5])) == 0
    lat.close()
