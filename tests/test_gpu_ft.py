"""Matsubara tail moments and iw <-> tau transforms: HIP path vs oracle and closed forms.

Transforms: 1e-12 normwise relative (both sides form the same FP64 phase
w*tau; only the sin/cos implementations differ).  Moments: the least-squares
normal equations are ill-conditioned (columns scale as 1, 1/w, 1/w^2, 1/w^3;
the condition number of the 2x2 (1/w, 1/w^3) block reaches ~1e9 over a
window), so a rounding difference in K^T K would move m2, m3 by ~1e-7.  Both
sides therefore run one operation sequence: the basis values from the same
std::pow, K^T K and its LU inverse with explicit fma, the rows of
KM_solve = (K^T K)^{-1} K^T formed before they meet g, as the reference
associates the product (FourierTransform.cpp:239-243), and no other
contraction.  The windows must be identical and the moments agree to 1e-12.
"""
import os

import numpy as np
import pytest

from dmftpack_amd import lattice as L
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PI = 3.1415926535897932384626


def rel(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="module")
def H():
    from dmftpack_amd import gpu as G
    h = G.Handle()
    yield h
    h.close()


def three_pole_g(d, nf, beta, seed):
    rng = np.random.default_rng(seed)
    z = 1j * PI * (2 * np.arange(nf) + 1) / beta
    fw = np.zeros((nf, d, d), complex)
    for _ in range(3):
        B = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        A = B @ B.conj().T / (3 * d)
        e = rng.uniform(-1.5, 1.5)
        fw += A[None] / (z[:, None, None] - e)
    return fw


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["poly", "pole5", "t2g", "pole10"])
def test_moments_match_oracle(H, kind):
    if kind == "poly":
        rng = np.random.default_rng(2)
        d, nf, beta = 3, 200, 30.0
        ms = []
        for s in (0.5, 1.0, 0.3, 0.2):
            X = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
            ms.append(s * (X + X.conj().T) / 2)
        ms[1] = ms[1] @ ms[1].conj().T + np.eye(d)
        z = 1j * PI * (2 * np.arange(nf) + 1) / beta
        fw = sum(ms[q][None] / z[:, None, None] ** q for q in range(4))
    elif kind == "pole5":
        fw, _ = L.pole_sigma(5, 300, 50.0, seed=3)
        beta = 50.0
    elif kind == "pole10":
        fw, _ = L.pole_sigma(10, 256, 100.0, seed=4, block=5)
        beta = 100.0
    else:
        fw, _ = L.t2g_sigma(256, 40.0, seed=1)
        beta = 40.0
    ref, wref = O.asymto_moments(fw, beta, return_windows=True)
    got, wgot = H.get_asymto_moments(fw, beta, return_windows=True)
    print(f"moments {kind}: windows GPU {wgot} oracle {wref}, rel err " +
          " ".join(f"{rel(got[q], ref[q]):.1e}" for q in range(4)), flush=True)
    assert wgot == wref
    for q in range(4):
        assert rel(got[q], ref[q]) <= 1e-12 or np.max(np.abs(got[q] - ref[q])) < 1e-15, (q, wgot, wref)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["all_rejected", "filtered"])
def test_moment_windows_constructed(H, case):
    """The constructed window cases of tests/test_cpu_oracle_choices.py: every window
    rejected by the positivity tests (the last centre for all moments), and the
    lowest-spread windows rejected with a clear admissible minimum elsewhere."""
    from moment_cases import noisy_two_pole
    fw = noisy_two_pole(200, 50.0, 1.2, 3.0, 1e-4, seed=1) if case == "all_rejected" else \
        noisy_two_pole(200, 50.0, 0.8, 1.0, 3e-4, seed=2)
    ref, wref = O.asymto_moments(fw, 50.0, return_windows=True)
    got, wgot = H.get_asymto_moments(fw, 50.0, return_windows=True)
    print(f"windows {case}: GPU {wgot} oracle {wref}", flush=True)
    assert wgot == wref
    for q in range(4):
        assert rel(got[q], ref[q]) <= 1e-12, q


@pytest.mark.gpu
def test_moments_cache_across_grids(H):
    """The handle keeps the fit's normal matrices for the last (N_freq, beta, branch):
    a sequence of calls that hits and misses it returns, bit for bit, what a
    fresh handle returns for each call."""
    from dmftpack_amd import gpu as G
    seq = [(5, 512, 50.0, False), (3, 512, 50.0, False), (5, 256, 50.0, False), (5, 512, 40.0, False),
           (4, 512, 40.0, True), (4, 512, 40.0, False), (5, 512, 50.0, False)]
    for d, nf, beta, green in seq:
        fw = three_pole_g(d, nf, beta, seed=d + nf)
        got, wg = H.get_asymto_moments(fw, beta, green_asym=green, return_windows=True)
        fresh = G.Handle()
        want, ww = fresh.get_asymto_moments(fw, beta, green_asym=green, return_windows=True)
        fresh.close()
        assert wg == ww and np.array_equal(got, want), (d, nf, beta, green)


@pytest.mark.gpu
def test_moments_green_branch(H):
    fw = three_pole_g(4, 300, 40.0, seed=5)
    ref, wref = O.asymto_moments(fw, 40.0, green_asym=True, return_windows=True)
    got, wgot = H.get_asymto_moments(fw, 40.0, green_asym=True, return_windows=True)
    assert wgot == wref
    assert np.array_equal(got[1], np.eye(4)) and np.max(np.abs(got[0])) == 0
    for q in (2, 3):
        assert rel(got[q], ref[q]) <= 1e-12, q


@pytest.mark.gpu
def test_ft_two_pole_closed_form(H):
    g = dict(np.load(os.path.join(GOLDEN, "ft_two_pole.npz")))
    gt, patched = H.fourier_transform(g["fw"], g["moments"], float(g["beta"]), int(g["n_tau"]))
    assert patched == 0
    assert rel(gt, g["gtau"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("d,nf,ntau,beta", [(1, 64, 50, 20.0), (5, 512, 1024, 50.0), (10, 256, 300, 100.0)])
def test_ft_w2tau_matches_oracle(H, d, nf, ntau, beta):
    fw = three_pole_g(d, nf, beta, seed=d)
    mom = O.asymto_moments(fw, beta, green_asym=True)
    ref, pr = O.ft_w2tau(fw, mom, beta, ntau)
    gt, pg = H.fourier_transform(fw, mom, beta, ntau)
    assert pg == pr
    assert rel(gt, ref) < 1e-12


@pytest.mark.gpu
def test_ft_causality_patch_matches_oracle(H):
    """A non-causal input (positive spectral weight) makes G(tau) positive somewhere."""
    beta, nf, ntau = 10.0, 64, 40
    z = 1j * PI * (2 * np.arange(nf) + 1) / beta
    fw = (1.0 / (z - 0.5) - 0.5 / (z + 1.0))[:, None, None]
    mom = np.zeros((4, 1, 1), complex)
    mom[1] = 1.0
    ref, pr = O.ft_w2tau(fw, mom, beta, ntau)
    gt, pg = H.fourier_transform(fw, mom, beta, ntau)
    assert pr == 1 and pg == 1
    assert rel(gt, ref) < 1e-12


@pytest.mark.gpu
def test_ft_tau2w_linear_and_oracle(H):
    g = dict(np.load(os.path.join(GOLDEN, "ft_linear_tau.npz")))
    fw, viol = H.ft_t_to_w(g["gtau"], float(g["beta"]), int(g["n_freq"]))
    assert rel(fw, g["fw"]) < 1e-12
    fref, vref = O.ft_tau2w(g["gtau"], float(g["beta"]), int(g["n_freq"]))
    assert viol == vref
    # a real G(tau) from the forward transform
    fw3 = three_pole_g(5, 128, 30.0, seed=9)
    mom = O.asymto_moments(fw3, 30.0, green_asym=True)
    gt, _ = O.ft_w2tau(fw3, mom, 30.0, 600)
    ref, vr = O.ft_tau2w(gt, 30.0, 128)
    got, vg = H.ft_t_to_w(gt, 30.0, 128)
    assert vg == vr
    assert rel(got, ref) < 1e-12
