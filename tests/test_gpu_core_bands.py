"""Core-band term of GiwToNele (GreenFtn.cpp:622-631) on the GPU path.

A downfolded lattice whose Kohn-Sham set has bands below and above the model
window (lattice.windowed_case).  With the bands bound (dmftgpu_set_core_bands)
N(mu), dN/dmu and the mu-search count every electron, as the reference's
NumberOfElectron does (example/SVO113/input.parm: N_ELECTRONS = 41, the total).
Checked against the oracle (itself pinned by the non-interacting count over all
KS bands, tests/test_cpu_core_bands.py) and against that closed form directly.
"""
import numpy as np
import pytest

from dmftpack_amd import lattice as L
from oracle import oracle as O

BETA = 9.0


def fermi(e, mu, beta):
    with np.errstate(over="ignore"):
        return 1.0 / (1.0 + np.exp(beta * (e - mu)))


@pytest.fixture(scope="module")
def gpu():
    import torch  # noqa: F401
    from dmftpack_amd import gpu as G
    return G


def case(padded=0, nc=3, nf=48, seed=7):
    w = L.windowed_case(6, 11, 4, seed=seed, n_below=2, padded=padded)
    sig, mom = L.pole_sigma(nc, nf, BETA, seed=seed + 1)
    return w, sig, O.asymto_moments(sig, BETA)


@pytest.mark.gpu
@pytest.mark.parametrize("padded", [0, 2])
def test_noninteracting_count_over_all_bands(gpu, padded):
    w, sig, _ = case(padded)
    nc, nf = sig.shape[1], sig.shape[0]
    zero = np.zeros_like(sig)
    lat = gpu.LatticeGF(w["Hk"], nc, nf, BETA, np.zeros((4, nc, nc), complex))
    lat.set_core_bands(w["ks_energy"], w["val_to_ks"])
    for method in (gpu.PLAIN_F64, gpu.SCHUR_F64, gpu.SCHUR_F32R):
        for mu in (-0.5, 0.2):
            f = fermi(w["ks_energy"], mu, BETA)
            exact, dexact = f.mean(0).sum(), (BETA * f * (1 - f)).mean(0).sum()
            N, dN = lat.density(zero, mu, method=method)
            print(f"padded {padded} method {method} mu {mu}: N {N:.15f} exact {exact:.15f}", flush=True)
            assert abs(N - exact) <= 1e-11 * exact
            assert abs(dN - dexact) <= 1e-10 * dexact
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("padded", [0, 2])
def test_density_matches_oracle_with_core_bands(gpu, padded):
    w, sig, mom = case(padded)
    core = (w["ks_energy"], w["val_to_ks"])
    lat = gpu.LatticeGF(w["Hk"], sig.shape[1], sig.shape[0], BETA, mom)
    lat.set_core_bands(*core)
    for mu in (-0.3, 0.25):
        N_ref, dN_ref = O.density(w["Hk"], sig, mom, mu, BETA, core=core)
        N_giw = O.giw_to_nele(w["Hk"], sig, mom, mu, BETA, core=core)
        assert abs(N_ref - N_giw) < 1e-12 * N_ref
        for method, tol in ((gpu.PLAIN_F64, 1e-12), (gpu.SCHUR_F64, 1e-12), (gpu.SCHUR_F32R, 1e-9)):
            N, dN = lat.density(sig, mu, method=method)
            out = lat.greenftn_w(sig, mu, method=method)
            print(f"padded {padded} method {method} mu {mu}: dN {abs(N - N_ref):.2e}", flush=True)
            assert abs(N - N_ref) <= tol * N_ref
            assert abs(dN - dN_ref) <= tol * dN_ref
            assert abs(out["n_total"] - N_ref) <= tol * N_ref
    # removing the bands gives the window's count back
    lat.set_core_bands(None, None)
    N0, _ = lat.density(sig, 0.1, method=gpu.PLAIN_F64)
    assert abs(N0 - O.density(w["Hk"], sig, mom, 0.1, BETA)[0]) <= 1e-12 * N0
    lat.close()


@pytest.mark.gpu
def test_adjust_mu_targets_total_count(gpu):
    w, sig, mom = case()
    core = (w["ks_energy"], w["val_to_ks"])
    lat = gpu.LatticeGF(w["Hk"], sig.shape[1], sig.shape[0], BETA, mom)
    lat.set_core_bands(*core)
    target = 5.2  # 2 bands below the window + 3.2 in it
    for method in (gpu.PLAIN_F64, gpu.SCHUR_F64):
        mu, nev = lat.adjust_mu(sig, 0.0, target, method=method, tol=1e-12)
        N_ref, dN_ref = O.density(w["Hk"], sig, mom, mu, BETA, core=core)
        print(f"method {method}: mu {mu:.12f} after {nev} evaluations, N - target {N_ref - target:.2e}", flush=True)
        assert abs(N_ref - target) <= 1e-9
        mu_ref, _ = O.adjust_mu(w["Hk"], sig, mom, 0.0, BETA, target, core=core)
        assert abs(mu - mu_ref) <= 2e-5
    lat.close()


@pytest.mark.gpu
def test_mu_bar_includes_core_derivative(gpu):
    """dL/dmu of L = ... + (N - N0)^2 carries 2 (N - N0) dN/dmu, core bands included."""
    w, sig, mom = case()
    core = (w["ks_energy"], w["val_to_ks"])
    lat = gpu.LatticeGF(w["Hk"], sig.shape[1], sig.shape[0], BETA, mom)
    gbar = np.zeros_like(sig)
    nbar, mu = 0.7, 0.1
    _, mb0 = lat.vjp(sig, mu, gbar, nbar, method=gpu.SCHUR_F64)
    lat.set_core_bands(*core)
    _, mb1 = lat.vjp(sig, mu, gbar, nbar, method=gpu.SCHUR_F64)
    _, dcore = O.core_bands(*core, mu, BETA)
    assert abs((mb1 - mb0) - nbar * dcore) <= 1e-12 * abs(nbar * dcore)
    lat.close()


@pytest.mark.gpu
def test_core_band_arguments_checked(gpu):
    from dmftpack_amd import capi
    w, sig, mom = case()
    lat = gpu.LatticeGF(w["Hk"], sig.shape[1], sig.shape[0], BETA, mom)
    bad = w["val_to_ks"].copy()
    bad[0, 0] = 11
    with pytest.raises(capi.DmftGpuError) as e:
        lat.set_core_bands(w["ks_energy"], bad)
    assert e.value.code == capi.ERR_ARG
    bad = w["val_to_ks"].copy()
    bad[1, 1] = bad[1, 0]
    with pytest.raises(capi.DmftGpuError) as e:
        lat.set_core_bands(w["ks_energy"], bad)
    assert e.value.code == capi.ERR_ARG
    lat.close()


@pytest.mark.gpu
def test_adjust_mu_rejects_unsupported_schur(gpu):
    """n_corr > 16 has no Schur kernels: every entry point, the mu-search included, says so."""
    from dmftpack_amd import capi
    Hk = L.random_tb_hk(20, 2, 3)
    sig, mom = L.pole_sigma(18, 24, 30.0, 4)
    lat = gpu.LatticeGF(Hk, 18, 24, 30.0, mom)
    for method in (gpu.SCHUR_F64, gpu.SCHUR_F32R):
        with pytest.raises(capi.DmftGpuError) as e:
            lat.adjust_mu(sig, 0.0, 9.0, method=method)
        assert e.value.code == capi.ERR_UNSUPPORTED
    with pytest.raises(capi.DmftGpuError) as e:
        lat.adjust_mu(sig, float("nan"), 9.0, method=gpu.PLAIN_F64)
    assert e.value.code == capi.ERR_ARG
    lat.close()
