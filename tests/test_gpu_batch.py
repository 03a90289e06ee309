"""Batched lattices (persistent per-n_c launches, hipGraph replay, autotuner)
against the single-lattice path on the same inputs.

Bar: SCHUR_F64 1e-12 normwise relative for every output of every system,
SCHUR_F32R 1e-9; every graph replay identical to the first (the step has no
atomics and a static work list, so replays are bit-for-bit deterministic).
"""
import numpy as np
import pytest

from dmftpack_amd import lattice as L


def rel(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / max(np.max(np.abs(ref)), 1e-300))


def systems(count=7, seed=20, nf=48):
    out = []
    shapes = [(3, 2), (5, 3), (7, 2), (10, 3), (14, 2), (20, 2), (6, 4)]
    for i in range(count):
        n, side = shapes[i % len(shapes)]
        nc = min(5, n)
        Hk = L.random_tb_hk(n, side, seed + i)
        sig, mom = L.pole_sigma(nc, nf, 40.0, seed + i)
        tgt = L.perturbed(sig, seed=seed + 100 + i)
        out.append(dict(Hk=Hk, sigma=sig, moments=mom, beta=40.0, mu=0.1 * (i % 3), tgt_sigma=tgt,
                        n_target=0.5 * nc))
    return out


@pytest.fixture(scope="module")
def G():
    from dmftpack_amd import gpu
    return gpu


def reference(G, sd, method):
    lat = G.LatticeGF(sd["Hk"], sd["sigma"].shape[1], sd["sigma"].shape[0], sd["beta"], sd["moments"])
    out = lat.greenftn_w(sd["sigma"], sd["mu"], method=method)
    gbar = 2.0 * (out["gloc"] - sd["gloc_target"])
    sb, mb = lat.vjp(sd["sigma"], sd["mu"], gbar, 2.0 * (out["n_total"] - sd["n_target"]), method=method)
    lat.close()
    out.update(sigma_bar=sb, mu_bar=mb)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("method,tol", [(1, 1e-12), (2, 1e-9)])
@pytest.mark.parametrize("graph,tune", [(True, True), (False, False)])
def test_batch_matches_single(G, method, tol, graph, tune):
    sys_ = systems()
    # targets: G_loc of the perturbed Sigma from the plain path
    for sd in sys_:
        lat = G.LatticeGF(sd["Hk"], sd["sigma"].shape[1], sd["sigma"].shape[0], sd["beta"], sd["moments"])
        sd["gloc_target"] = lat.greenftn_w(sd["tgt_sigma"], sd["mu"], method=G.PLAIN_F64)["gloc"]
        lat.close()
    h = G.Handle()
    b = G.Batch(h, sys_, method=method, tune=tune, graph=graph, prepare=True)
    print(b.tuning(), flush=True)
    runs = []
    for rep in range(3):
        b.run()
        runs.append(b.results())
        _ = G.torch.ones(1000, device="cuda").sum().item()  # eager work between replays
    for i, sd in enumerate(sys_):
        ref = reference(G, sd, method)
        got = runs[0][i]
        assert rel(got["gloc"], ref["gloc"]) <= tol, i
        assert rel(got["num_matrix"], ref["num_matrix"]) <= tol, i
        assert abs(got["n_total"] - ref["n_total"]) <= tol * max(1.0, abs(ref["n_total"])), i
        assert rel(got["sigma_bar"], ref["sigma_bar"]) <= max(tol, 1e-9), i
        assert abs(got["mu_bar"] - ref["mu_bar"]) <= max(tol, 1e-9) * max(1.0, abs(ref["mu_bar"])), i
        for r in runs[1:]:
            for key in ("gloc", "num_matrix", "sigma_bar"):
                assert np.array_equal(r[i][key], got[key]), (i, key)
            assert r[i]["n_total"] == got["n_total"] and r[i]["mu_bar"] == got["mu_bar"]
    b.close()
    h.close()


def systems_wide(seed=40, nf=40):
    """Mixed n_corr groups, including column-per-lane ones (n_corr 8 and 10,
    two systems each) next to a one-matrix-per-lane group."""
    out = []
    for i, (n, nc, side) in enumerate([(12, 8, 2), (16, 8, 3), (20, 10, 2), (14, 10, 2), (9, 5, 3)]):
        Hk = L.random_tb_hk(n, side, seed + i)
        sig, mom = L.pole_sigma(nc, nf, 40.0, seed + i)
        tgt = L.perturbed(sig, seed=seed + 100 + i)
        out.append(dict(Hk=Hk, sigma=sig, moments=mom, beta=40.0, mu=0.05 * i, tgt_sigma=tgt, n_target=0.5 * nc))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("method,tol", [(1, 1e-12), (2, 1e-9)])
def test_batch_wide_groups_match_single(G, method, tol):
    """n_corr > 6 runs in the batch as persistent column-per-lane launches; every
    output of every system equals its single-lattice run, replays are identical."""
    sys_ = systems_wide()
    for sd in sys_:
        lat = G.LatticeGF(sd["Hk"], sd["sigma"].shape[1], sd["sigma"].shape[0], sd["beta"], sd["moments"])
        sd["gloc_target"] = lat.greenftn_w(sd["tgt_sigma"], sd["mu"], method=G.PLAIN_F64)["gloc"]
        lat.close()
    h = G.Handle()
    b = G.Batch(h, sys_, method=method, tune=True, graph=True, prepare=True)
    print(b.tuning(), flush=True)
    runs = []
    for rep in range(2):
        b.run()
        runs.append(b.results())
        _ = G.torch.ones(1000, device="cuda").sum().item()  # eager work between replays
    for i, sd in enumerate(sys_):
        ref = reference(G, sd, method)
        got = runs[0][i]
        print(f"system {i}: n={sd['Hk'].shape[1]} nc={sd['sigma'].shape[1]}", flush=True)
        assert rel(got["gloc"], ref["gloc"]) <= tol, i
        assert rel(got["num_matrix"], ref["num_matrix"]) <= tol, i
        assert abs(got["n_total"] - ref["n_total"]) <= tol * max(1.0, abs(ref["n_total"])), i
        assert rel(got["sigma_bar"], ref["sigma_bar"]) <= max(tol, 1e-9), i
        assert abs(got["mu_bar"] - ref["mu_bar"]) <= max(tol, 1e-9) * max(1.0, abs(ref["mu_bar"])), i
        for key in ("gloc", "num_matrix", "sigma_bar"):
            assert np.array_equal(runs[1][i][key], got[key]), (i, key)
    b.close()
    h.close()


@pytest.mark.gpu
def test_batch_rejects_unsupported(G):
    sd = systems(1)[0]
    sd["sigma"] = np.zeros((16, 17, 17), complex)  # n_corr = 17 > 16
    sd["moments"] = np.zeros((4, 17, 17), complex)
    sd["Hk"] = L.random_tb_hk(20, 2, 1)
    h = G.Handle()
    with pytest.raises(G.capi.DmftGpuError):
        G.Batch(h, [sd], method=1)
    h.close()
