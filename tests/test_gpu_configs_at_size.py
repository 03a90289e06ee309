"""Every benchmark configuration's workload at its full size on the GPU path.

- Batching (64 systems, n in {3,...,20}, Nk in {8^3,...,24^3}, N_freq = 1024,
  beta = 40): the batched loss + gradient step (FP64 and FP32+refine, hipGraph
  replay) against each system's single-lattice plain FP64 run (the reference's
  full inverse), and a batch of every system's k subset against the oracle at
  all frequencies (forward and vector-Jacobian product).
- Large (n = 32, two 5x5 clusters, Nk = 40^3, N_freq = 4096, beta = 100): the
  adjoint on the full mesh -- FP32+refine against FP64, and central finite
  differences (h = 1e-6) of the GPU loss in Sigma entries and mu -- and the
  oracle's vector-Jacobian product on a 4^3 mesh at all 4096 frequencies;
  the iw -> tau transform on N_tau = 8192 (d = 10) against the oracle's
  FourierTransform_tau at 48 tau points, and tau -> iw back against FT_t_to_w.
- t2g (Nk = 16^3, N_freq = 1024, beta = 40): G_loc(tau) on N_tau = 2048 from the
  GPU G_loc against the oracle's whole transform, and G(0) + G(beta) = -1.
"""
import numpy as np
import pytest

from dmftpack_amd import lattice as L
from oracle import oracle as O

NTHR = 16


def rel(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="module")
def G():
    from dmftpack_amd import gpu
    return gpu


@pytest.fixture(scope="module")
def batch64():
    sys_ = L.batch_systems()
    for i, sd in enumerate(sys_):
        sd["moments"] = sd["exact_moments"]
        sd["mu"] = 0.1 * ((i % 5) - 2)
        sd["n_target"] = 0.45 * sd["n"]
        sd["tgt_sigma"] = L.perturbed(sd["sigma"], seed=900 + i)
    return sys_


def plain_reference(G, sd):
    lat = G.LatticeGF(sd["Hk"], sd["sigma"].shape[1], sd["sigma"].shape[0], sd["beta"], sd["moments"])
    if "gloc_target" not in sd:
        sd["gloc_target"] = lat.greenftn_w(sd["tgt_sigma"], sd["mu"], method=G.PLAIN_F64)["gloc"]
    out = lat.greenftn_w(sd["sigma"], sd["mu"], method=G.PLAIN_F64)
    gbar = 2.0 * (out["gloc"] - sd["gloc_target"])
    sb, mb = lat.vjp(sd["sigma"], sd["mu"], gbar, 2.0 * (out["n_total"] - sd["n_target"]), method=G.PLAIN_F64)
    lat.close()
    out.update(sigma_bar=sb, mu_bar=mb)
    return out


@pytest.mark.gpu
def test_batch64_against_single_lattice_plain(G, batch64):
    n_units = sum(sd["Hk"].shape[0] * 2 * sd["sigma"].shape[0] for sd in batch64)
    print(f"batch64: {len(batch64)} systems, {n_units:.3e} (k, w) units", flush=True)
    refs = [plain_reference(G, sd) for sd in batch64]
    print("batch64: plain FP64 references done", flush=True)
    h = G.Handle()
    for method, tol in ((G.SCHUR_F64, 1e-12), (G.SCHUR_F32R, 1e-9)):
        b = G.Batch(h, batch64, method=method, tune=True, graph=True)
        b.run()
        first = b.results()
        b.run()
        again = b.results()
        worst = dict(g=0.0, n=0.0, s=0.0, m=0.0)
        for i, (got, ref) in enumerate(zip(first, refs)):
            eg = rel(got["gloc"], ref["gloc"])
            en = abs(got["n_total"] - ref["n_total"]) / max(1.0, abs(ref["n_total"]))
            es = rel(got["sigma_bar"], ref["sigma_bar"])
            em = abs(got["mu_bar"] - ref["mu_bar"]) / max(1.0, abs(ref["mu_bar"]))
            worst = dict(g=max(worst["g"], eg), n=max(worst["n"], en), s=max(worst["s"], es), m=max(worst["m"], em))
            assert eg <= tol and rel(got["num_matrix"], ref["num_matrix"]) <= tol and en <= tol, (i, method)
            assert es <= max(tol, 1e-9) and em <= max(tol, 1e-9), (i, method, es, em)
            assert np.array_equal(again[i]["gloc"], got["gloc"]) and np.array_equal(again[i]["sigma_bar"],
                                                                                   got["sigma_bar"])
        print(f"batch64 method {method}: worst rel G_loc {worst['g']:.1e} N {worst['n']:.1e} "
              f"sigma_bar {worst['s']:.1e} mu_bar {worst['m']:.1e}", flush=True)
        b.close()
    h.close()


@pytest.mark.gpu
def test_batch64_k_subsets_against_oracle(G, batch64):
    """The batch over every system's k subset (every 64th k point, at least 8), all
    1024 frequencies, against the oracle's GreenFtn_w and vector-Jacobian product."""
    subs = []
    for i, sd in enumerate(batch64):
        Hk = sd["Hk"][::64] if sd["Hk"].shape[0] >= 512 else sd["Hk"][::sd["Hk"].shape[0] // 8]
        s = dict(sd, Hk=Hk)
        s.pop("gloc_target", None)
        lat = G.LatticeGF(Hk, s["sigma"].shape[1], s["sigma"].shape[0], s["beta"], s["moments"])
        s["gloc_target"] = lat.greenftn_w(s["tgt_sigma"], s["mu"], method=G.PLAIN_F64)["gloc"]
        lat.close()
        subs.append(s)
    h = G.Handle()
    b = G.Batch(h, subs, method=G.SCHUR_F64, tune=False, graph=True)
    b.run()
    res = b.results()
    worst = 0.0
    for i, (s, got) in enumerate(zip(subs, res)):
        ref = O.greenftn_w(s["Hk"], s["sigma"], s["moments"], s["mu"], s["beta"], nthreads=NTHR)
        N_ref, _ = O.density(s["Hk"], s["sigma"], s["moments"], s["mu"], s["beta"], nthreads=NTHR)
        gbar = 2.0 * (got["gloc"] - s["gloc_target"])
        sb, mb = O.gloc_vjp(s["Hk"], s["sigma"], s["moments"], s["mu"], s["beta"], gbar,
                            2.0 * (got["n_total"] - s["n_target"]), nthreads=NTHR)
        e = max(rel(got["gloc"], ref["gloc"]), rel(got["num_matrix"], ref["num_matrix"]),
                abs(got["n_total"] - N_ref) / N_ref, rel(got["sigma_bar"], sb),
                abs(got["mu_bar"] - mb) / max(1.0, abs(mb)))
        worst = max(worst, e)
        assert e <= 1e-12, (i, e)
    print(f"batch64 k subsets vs oracle: worst {worst:.1e}", flush=True)
    b.close()
    h.close()


@pytest.mark.gpu
def test_large32_adjoint_full_mesh(G):
    w = L.workload("large32", seed=1)
    Hk, sig, beta, mom = w["Hk"], w["sigma"], w["beta"], w["exact_moments"]
    nf, mu, N0 = sig.shape[0], 0.0, w["n_target"]
    print(f"large32 adjoint full mesh: nk={Hk.shape[0]} nf={nf}", flush=True)
    lat = G.LatticeGF(Hk, 10, nf, beta, mom, n_cluster=2)
    del w, Hk
    tgt = lat.greenftn_w(L.perturbed(sig, seed=7), mu, method=G.SCHUR_F64)["gloc"]

    def loss(s, m):
        o = lat.greenftn_w(s, m, method=G.SCHUR_F64)
        return float(np.sum(np.abs(o["gloc"] - tgt) ** 2) + (o["n_total"] - N0) ** 2), o

    L0, o = loss(sig, mu)
    gbar = 2.0 * (o["gloc"] - tgt)
    nbar = 2.0 * (o["n_total"] - N0)
    sb64, mb64 = lat.vjp(sig, mu, gbar, nbar, method=G.SCHUR_F64)
    sb32, mb32 = lat.vjp(sig, mu, gbar, nbar, method=G.SCHUR_F32R)
    print(f"large32: F32R vs F64 rel sigma_bar {rel(sb32, sb64):.1e}, mu_bar {abs(mb32 - mb64):.1e}", flush=True)
    assert rel(sb32, sb64) <= 1e-8 and abs(mb32 - mb64) <= 1e-8 * max(1.0, abs(mb64))
    # loss + gradient in one call agrees with the two-call form
    ll = lat.loss_grad(sig, mu, tgt, N0, method=G.SCHUR_F64)
    assert abs(ll["loss"] - L0) <= 1e-12 * L0 and rel(ll["sigma_bar"], sb64) <= 1e-12
    # h = 1e-4: the loss sums 4096 x 100 entries, so its rounding (~1e-15 of L) over
    # h = 1e-6 would reach the 1e-6 bar on the small entries; the central difference's
    # truncation error is O(h^2)
    h = 1e-4
    # intra-cluster entries at low frequency (the largest gradients) and one high one
    for n, a, b in ((0, 0, 0), (3, 6, 8), (nf // 3, 2, 4)):
        for unit in (1.0, 1j):
            sp, sm = sig.copy(), sig.copy()
            sp[n, a, b] += h * unit
            sm[n, a, b] -= h * unit
            fd = (loss(sp, mu)[0] - loss(sm, mu)[0]) / (2 * h)
            an = sb64[n, a, b].real if unit == 1.0 else sb64[n, a, b].imag
            print(f"large32 FD ({n},{a},{b},{'re' if unit == 1.0 else 'im'}): {an:.9e} vs {fd:.9e}", flush=True)
            assert abs(an - fd) <= 1e-6 * max(abs(fd), np.max(np.abs(sb64))), (n, a, b, an, fd)
    fd_mu = (loss(sig, mu + h)[0] - loss(sig, mu - h)[0]) / (2 * h)
    assert abs(mb64 - fd_mu) <= 1e-6 * max(1.0, abs(fd_mu))
    # a random direction over every entry: dL = Re sum conj(sigma_bar) dSigma + mu_bar dmu
    rng = np.random.default_rng(11)
    dsig = (rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)) * 1e-2
    dsig[:, :5, 5:] = 0
    dsig[:, 5:, :5] = 0
    dmu = 0.3
    an = float(np.sum((sb64.conj() * dsig).real) + mb64 * dmu)
    fd = (loss(sig + h * dsig, mu + h * dmu)[0] - loss(sig - h * dsig, mu - h * dmu)[0]) / (2 * h)
    print(f"large32 FD along a random direction: {an:.12e} vs {fd:.12e}", flush=True)
    assert abs(an - fd) <= 1e-6 * abs(fd)
    lat.close()


@pytest.mark.gpu
def test_large32_adjoint_oracle_all_frequencies(G):
    w = L.workload("large32", seed=1, nk_side=4)
    Hk, sig, beta, mom = w["Hk"], w["sigma"], w["beta"], w["exact_moments"]
    nf = sig.shape[0]
    rng = np.random.default_rng(5)
    gbar = rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)
    print(f"large32 oracle VJP: nk={Hk.shape[0]} nf={nf}", flush=True)
    sb_ref, mb_ref = O.gloc_vjp(Hk, sig, mom, 0.05, beta, gbar, 0.4, nthreads=NTHR)
    lat = G.LatticeGF(Hk, 10, nf, beta, mom, n_cluster=2)
    for method, tol in ((G.SCHUR_F64, 1e-12), (G.SCHUR_F32R, 1e-8), (G.PLAIN_F64, 1e-12)):
        sb, mb = lat.vjp(sig, 0.05, gbar, 0.4, method=method)
        print(f"large32 VJP method {method}: rel {rel(sb, sb_ref):.1e}", flush=True)
        assert rel(sb, sb_ref) <= tol and abs(mb - mb_ref) <= tol * max(1.0, abs(mb_ref)), method
    lat.close()


def three_pole_g(d, nf, beta, seed):
    rng = np.random.default_rng(seed)
    z = 1j * np.pi * (2 * np.arange(nf) + 1) / beta
    fw = np.zeros((nf, d, d), complex)
    for wgt in (0.5, 0.3, 0.2):
        B = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        q, _ = np.linalg.qr(B)
        e = rng.uniform(-1.5, 1.5, size=d)
        fw += wgt * np.einsum("ab,wb,cb->wac", q, 1.0 / (z[:, None] - e[None, :]), q.conj())
    return fw


@pytest.mark.gpu
def test_large32_transforms_full_grid(G):
    d, nf, ntau, beta = 10, 4096, 8192, 100.0
    fw = three_pole_g(d, nf, beta, seed=8)
    H = G.Handle()
    mom = H.get_asymto_moments(fw, beta, green_asym=True)
    mref = O.asymto_moments(fw, beta, green_asym=True)
    assert rel(mom[2:], mref[2:]) <= 1e-12
    gt, patched = H.fourier_transform(fw, mom, beta, ntau)
    taus = np.linspace(0, ntau, 48).astype(int)
    worst = 0.0
    for j in taus:
        ref = O.ft_tau(fw, mom, beta, j * beta / ntau)
        worst = max(worst, rel(gt[j], ref))
    print(f"iw->tau d=10 N_tau=8192: worst rel {worst:.1e} over {len(taus)} tau points, patched {patched}", flush=True)
    assert worst <= 1e-12 and patched == 0
    assert np.max(np.abs(gt[0] + gt[-1] + np.eye(d))) < 1e-10  # G(0) + G(beta) = -m1 = -1
    back, viol = H.ft_t_to_w(gt, beta, nf)
    ref_back = O.ft_tau2w(gt, beta, nf)
    print(f"tau->iw d=10: rel {rel(back, ref_back[0]):.1e}, violations {viol}/{ref_back[1]}", flush=True)
    assert rel(back, ref_back[0]) <= 1e-12 and viol == ref_back[1]
    H.close()


@pytest.mark.gpu
def test_t2g_gtau_full_grid(G):
    w = L.workload("t2g", seed=1)
    Hk, sig, beta = w["Hk"], w["sigma"], w["beta"]
    mom = O.asymto_moments(sig, beta)
    lat = G.LatticeGF(Hk, 3, sig.shape[0], beta, mom)
    gl = lat.greenftn_w(sig, 0.1, method=G.SCHUR_F64)["gloc"]
    gm = lat.get_asymto_moments(gl, beta, green_asym=True)
    gt, patched = lat.fourier_transform(gl, gm, beta, w["n_tau"])
    gm_ref = O.asymto_moments(gl, beta, green_asym=True)
    gt_ref, patched_ref = O.ft_w2tau(gl, gm_ref, beta, w["n_tau"], nthreads=NTHR)
    print(f"t2g G(tau) N_tau=2048: rel {rel(gt, gt_ref):.1e}, patched {patched}/{patched_ref}", flush=True)
    assert rel(gt, gt_ref) <= 1e-12 and patched == patched_ref
    assert np.max(np.abs(gt[0] + gt[-1] + np.eye(3))) < 1e-8
    lat.close()
