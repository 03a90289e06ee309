"""Schur-complement paths against the FP64 oracle.

n_c <= 6 runs the one-matrix-per-lane kernels, 7 <= n_c <= 16 the
column-per-lane kernels (buckets 8/10/12/16, with padding), n_c = n the
degenerate Schur complement with no d block.

SCHUR_F64: same bar as the plain path, 1e-12 normwise relative.
SCHUR_F32R (FP32 solve + FP64-residual refinement): max |dG_loc| <= 1e-9 and
relative 1e-9 on NumMatrix, N, dN/dmu; gradients relative 1e-6.
"""
import numpy as np
import pytest

from dmftpack_amd import lattice as L
from oracle import oracle as O


def rel(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


def case(name, seed=2):
    if name == "dp14":
        w = L.workload("dp14", seed=seed, nk_side=4, n_freq=40)
        return w["Hk"], w["sigma"], w["beta"], 0.3
    n, nc = (int(v) for v in name.split("_")[1:])
    Hk = L.random_tb_hk(n, 3, seed)
    sig, _ = L.pole_sigma(nc, 36, 30.0, seed)
    return Hk, sig, 30.0, 0.2


CASES = ["dp14", "c_6_3", "c_9_4", "c_20_5", "c_8_6", "c_4_1", "c_3_3", "c_5_2",
         "c_12_7", "c_16_8", "c_16_10", "c_11_9", "c_10_10", "c_20_12", "c_24_16", "c_16_16"]
METHODS = [("f64", 1, 1e-12), ("f32r", 2, 1e-9)]


@pytest.fixture(scope="module")
def gpu():
    from dmftpack_amd import gpu as G
    return G


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("mname,method,tol", METHODS)
def test_greenftn_w_schur(gpu, name, mname, method, tol):
    Hk, sig, beta, mu = case(name)
    mom = O.asymto_moments(sig, beta)
    print(f"case {name} {mname}: n={Hk.shape[1]} nc={sig.shape[1]}", flush=True)
    ref = O.greenftn_w(Hk, sig, mom, mu, beta)
    lat = gpu.LatticeGF(Hk, sig.shape[1], sig.shape[0], beta, mom)
    out = lat.greenftn_w(sig, mu, method=method)
    assert np.max(np.abs(out["gloc"] - ref["gloc"])) <= tol
    assert rel(out["gloc"], ref["gloc"]) <= tol
    assert rel(out["num_matrix"], ref["num_matrix"]) <= tol
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["dp14", "c_9_4", "c_3_3", "c_20_5", "c_12_7", "c_16_10", "c_10_10", "c_24_16"])
@pytest.mark.parametrize("mname,method,tol", METHODS)
def test_density_schur(gpu, name, mname, method, tol):
    Hk, sig, beta, mu = case(name)
    mom = O.asymto_moments(sig, beta)
    N_ref, dN_ref = O.density(Hk, sig, mom, mu, beta)
    lat = gpu.LatticeGF(Hk, sig.shape[1], sig.shape[0], beta, mom)
    N, dN = lat.density(sig, mu, method=method)
    assert abs(N - N_ref) <= tol * max(abs(N_ref), 1.0)
    assert abs(dN - dN_ref) <= tol * max(abs(dN_ref), 1.0)
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["dp14", "c_9_4", "c_20_5", "c_12_7", "c_16_10", "c_11_9", "c_20_12", "c_24_16"])
@pytest.mark.parametrize("mname,method,tol", [("f64", 1, 1e-12), ("f32r", 2, 1e-8)])
def test_vjp_schur(gpu, name, mname, method, tol):
    Hk, sig, beta, mu = case(name)
    mom = O.asymto_moments(sig, beta)
    rng = np.random.default_rng(5)
    gbar = rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)
    nbar = -0.4
    sb_ref, mb_ref = O.gloc_vjp(Hk, sig, mom, mu, beta, gbar, nbar)
    lat = gpu.LatticeGF(Hk, sig.shape[1], sig.shape[0], beta, mom)
    sb, mb = lat.vjp(sig, mu, gbar, nbar, method=method)
    assert rel(sb, sb_ref) <= tol
    assert abs(mb - mb_ref) <= tol * max(abs(mb_ref), 1.0)
    lat.close()


@pytest.mark.gpu
def test_schur_matches_plain_at_headline_shape(gpu):
    """n=14, n_c=5 at a mid-size mesh: Schur F64 vs plain F64, F32R vs plain."""
    w = L.workload("dp14", seed=3, nk_side=8, n_freq=256)
    Hk, sig, beta = w["Hk"], w["sigma"], w["beta"]
    lat = gpu.LatticeGF(Hk, 5, sig.shape[0], beta, w["exact_moments"])
    p = lat.greenftn_w(sig, 0.25, method=gpu.PLAIN_F64)
    s = lat.greenftn_w(sig, 0.25, method=gpu.SCHUR_F64)
    r = lat.greenftn_w(sig, 0.25, method=gpu.SCHUR_F32R)
    assert rel(s["gloc"], p["gloc"]) <= 1e-12
    assert np.max(np.abs(r["gloc"] - p["gloc"])) <= 1e-9
    assert rel(r["num_matrix"], p["num_matrix"]) <= 1e-9
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mname,method,tol", METHODS)
def test_large32_two_clusters(gpu, mname, method, tol):
    """The 32-orbital shape: two 5x5 correlated blocks (n_c = 10, NumCluster = 2,
    G_loc returned block-diagonal), small mesh."""
    w = L.workload("large32", seed=4, nk_side=2, n_freq=40)
    Hk, sig, beta = w["Hk"], w["sigma"], w["beta"]
    mom = O.asymto_moments(sig, beta)
    ref = O.greenftn_w(Hk, sig, mom, -0.2, beta, n_cluster=2)
    lat = gpu.LatticeGF(Hk, 10, sig.shape[0], beta, mom, n_cluster=2)
    out = lat.greenftn_w(sig, -0.2, method=method)
    assert np.max(np.abs(out["gloc"] - ref["gloc"])) <= tol
    assert rel(out["num_matrix"], ref["num_matrix"]) <= tol
    N_ref, dN_ref = O.density(Hk, sig, mom, -0.2, beta)
    N, dN = lat.density(sig, -0.2, method=method)
    assert abs(N - N_ref) <= tol * max(abs(N_ref), 1.0)
    assert abs(dN - dN_ref) <= tol * max(abs(dN_ref), 1.0)
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["dp14", "c_6_3", "c_4_1", "c_8_6", "c_5_2", "c_3_3"])
@pytest.mark.parametrize("mname,method,tol", METHODS)
def test_column_layout_small_nc(gpu, name, mname, method, tol):
    """n_c <= 6 through the column-per-lane kernels (buckets 4/5/6): forward,
    density and adjoint against the oracle."""
    Hk, sig, beta, mu = case(name)
    mom = O.asymto_moments(sig, beta)
    print(f"columns {name} {mname}", flush=True)
    lat = gpu.LatticeGF(Hk, sig.shape[1], sig.shape[0], beta, mom)
    lat.set_option(gpu.capi.OPT_SCHUR_LAYOUT, 1)
    ref = O.greenftn_w(Hk, sig, mom, mu, beta)
    out = lat.greenftn_w(sig, mu, method=method)
    assert np.max(np.abs(out["gloc"] - ref["gloc"])) <= tol
    assert rel(out["num_matrix"], ref["num_matrix"]) <= tol
    N_ref, dN_ref = O.density(Hk, sig, mom, mu, beta)
    N, dN = lat.density(sig, mu, method=method)
    assert abs(N - N_ref) <= tol * max(abs(N_ref), 1.0)
    assert abs(dN - dN_ref) <= tol * max(abs(dN_ref), 1.0)
    rng = np.random.default_rng(5)
    gbar = rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)
    sb_ref, mb_ref = O.gloc_vjp(Hk, sig, mom, mu, beta, gbar, -0.4)
    sb, mb = lat.vjp(sig, mu, gbar, -0.4, method=method)
    gtol = 1e-12 if method == 1 else 1e-8
    assert rel(sb, sb_ref) <= gtol
    assert abs(mb - mb_ref) <= gtol * max(abs(mb_ref), 1.0)
    lat.close()


@pytest.mark.gpu
def test_update_moments_matches_prepare(gpu):
    """New tail moments through dmftgpu_update_moments give what a fresh
    prepare gives (the Schur data of H(k) is reused)."""
    Hk, sig, beta, mu = case("dp14")
    mom = O.asymto_moments(sig, beta)
    sig2 = L.perturbed(sig, seed=9)
    mom2 = O.asymto_moments(sig2, beta)
    lat = gpu.LatticeGF(Hk, 5, sig.shape[0], beta, mom)
    lat.greenftn_w(sig, mu, method=gpu.SCHUR_F64)
    lat.set_moments(mom2)
    a = lat.greenftn_w(sig2, mu, method=gpu.SCHUR_F64)
    fresh = gpu.LatticeGF(Hk, 5, sig.shape[0], beta, mom2)
    b = fresh.greenftn_w(sig2, mu, method=gpu.SCHUR_F64)
    ref = O.greenftn_w(Hk, sig2, mom2, mu, beta)
    assert np.array_equal(a["gloc"], b["gloc"]) and np.array_equal(a["num_matrix"], b["num_matrix"])
    assert rel(a["num_matrix"], ref["num_matrix"]) <= 1e-12
    lat.close()
    fresh.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,ncl", [("dp14", 1), ("c_9_4", 1), ("c_4_1", 1), ("c_8_6", 1), ("c_16_10", 1),
                                      ("large32", 2)])
@pytest.mark.parametrize("mname,method,tol", [("plain", 0, 1e-12), ("f64", 1, 1e-12), ("f32r", 2, 1e-8)])
def test_loss_grad(gpu, name, ncl, mname, method, tol):
    """dmftgpu_gloc_loss_grad (fused on the one-matrix-per-lane path) equals the
    forward, the cotangents and the adjoint called one by one, and the oracle."""
    if name == "large32":
        w = L.workload("large32", seed=4, nk_side=2, n_freq=24)
        Hk, sig, beta, mu = w["Hk"], w["sigma"], w["beta"], -0.2
    else:
        Hk, sig, beta, mu = case(name)
    nc = sig.shape[1]
    mom = O.asymto_moments(sig, beta)
    print(f"loss_grad {name} {mname}", flush=True)
    lat = gpu.LatticeGF(Hk, nc, sig.shape[0], beta, mom, n_cluster=ncl)
    tgt = lat.greenftn_w(L.perturbed(sig, seed=11), mu, method=gpu.PLAIN_F64)["gloc"]
    n0 = 0.4 * nc
    r = lat.loss_grad(sig, mu, tgt, n0, method=method)
    fwd = lat.greenftn_w(sig, mu, method=method)
    assert np.array_equal(r["gloc"], fwd["gloc"]) and np.array_equal(r["num_matrix"], fwd["num_matrix"])
    N, _ = lat.density(sig, mu, method=method)
    assert abs(r["n_total"] - N) <= 1e-12 * max(1.0, abs(N))
    loss_ref = np.sum(np.abs(fwd["gloc"] - tgt) ** 2) + (r["n_total"] - n0) ** 2
    assert abs(r["loss"] - loss_ref) <= 1e-12 * loss_ref
    # the cotangent of the inter-cluster entries is zero (G_loc is returned block-diagonal)
    npc = nc // ncl
    gbar = 2.0 * (fwd["gloc"] - tgt)
    for a in range(nc):
        for b in range(nc):
            if a // npc != b // npc:
                gbar[:, a, b] = 0.0
    sb, mb = lat.vjp(sig, mu, gbar, 2.0 * (r["n_total"] - n0), method=method)
    assert rel(r["sigma_bar"], sb) <= 1e-12
    assert abs(r["mu_bar"] - mb) <= 1e-12 * max(1.0, abs(mb))
    sb_ref, mb_ref = O.gloc_vjp(Hk, sig, mom, mu, beta, gbar, 2.0 * (r["n_total"] - n0))
    assert rel(r["sigma_bar"], sb_ref) <= tol
    assert abs(r["mu_bar"] - mb_ref) <= tol * max(1.0, abs(mb_ref))
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("method", [1, 2])
def test_k_chunking_does_not_change_results(gpu, method):
    """The k chunking (DMFTGPU_OPT_TARGET_WORKGROUPS: 0 = default, 7 = one chunk,
    10^6 = one k per chunk) changes only the order of the k sums: the fused loss
    and gradient agree to rounding."""
    w = L.workload("dp14", seed=5, nk_side=5, n_freq=96)
    Hk, sig, beta = w["Hk"], w["sigma"], w["beta"]
    mom = w["exact_moments"]
    lat = gpu.LatticeGF(Hk, 5, 96, beta, mom)
    tgt = lat.greenftn_w(L.perturbed(sig, seed=3), 0.2, method=gpu.PLAIN_F64)["gloc"]
    out = []
    for target in (0, 7, 1000000):
        lat.set_option(gpu.capi.OPT_TARGET_WORKGROUPS, target)
        out.append(lat.loss_grad(sig, 0.2, tgt, 6.5, method=method))
    lat.set_option(gpu.capi.OPT_TARGET_WORKGROUPS, 0)
    lat.close()
    for r in out[1:]:
        assert rel(r["gloc"], out[0]["gloc"]) <= 1e-13
        assert rel(r["num_matrix"], out[0]["num_matrix"]) <= 1e-13
        assert abs(r["n_total"] - out[0]["n_total"]) <= 1e-12
        assert rel(r["sigma_bar"], out[0]["sigma_bar"]) <= 1e-12
        assert abs(r["mu_bar"] - out[0]["mu_bar"]) <= 1e-12 * max(1.0, abs(out[0]["mu_bar"]))


@pytest.mark.gpu
@pytest.mark.parametrize("target", [0, 7, 126])
def test_kpair_forward_matches_one_wavefront(gpu, target):
    """The FP32R forward over the first N_freq frequencies at n_c >= 5 runs by
    default as two wavefronts splitting each k chunk (every other k) into one set
    of LDS sums, added in turn; an explicit width of 64 runs it as one wavefront.
    The G_cc sums are then taken in the same k order, so G_loc and NumMatrix agree
    to rounding and repeat exactly; N differs only in the order of its tr G sums.
    Targets: default chunking, one chunk of 125 k (odd), 3 k per chunk (odd, and
    a last chunk of 2)."""
    w = L.workload("dp14", seed=5, nk_side=5, n_freq=96)
    Hk, sig, beta = w["Hk"], w["sigma"], w["beta"]
    lat = gpu.LatticeGF(Hk, 5, 96, beta, w["exact_moments"])
    lat.set_option(gpu.capi.OPT_TARGET_WORKGROUPS, target)
    split = lat.greenftn_w(sig, 0.2, method=gpu.SCHUR_F32R)
    again = lat.greenftn_w(sig, 0.2, method=gpu.SCHUR_F32R)
    lat.set_option(gpu.capi.OPT_SCHUR_WG, 64)
    one = lat.greenftn_w(sig, 0.2, method=gpu.SCHUR_F32R)
    lat.set_option(gpu.capi.OPT_SCHUR_WG, 0)
    ref = lat.greenftn_w(sig, 0.2, method=gpu.SCHUR_F64)
    lat.close()
    print(f"target {target}: G_loc bitwise equal to one wavefront: {np.array_equal(split['gloc'], one['gloc'])}",
          flush=True)
    assert np.array_equal(split["gloc"], again["gloc"]) and split["n_total"] == again["n_total"]
    assert rel(split["gloc"], one["gloc"]) <= 1e-14
    assert rel(split["num_matrix"], one["num_matrix"]) <= 1e-14
    assert abs(split["n_total"] - one["n_total"]) <= 1e-13 * max(1.0, abs(one["n_total"]))
    assert np.max(np.abs(split["gloc"] - ref["gloc"])) <= 1e-9
