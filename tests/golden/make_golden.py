"""Writes the golden fixtures of tests/golden/*.npz.

Expected outputs come from an independent numpy brute force of the reference
formulas (not from the oracle):
  GreenFtn_w (GreenFtn.cpp:142-422): numpy.linalg.inv per (k, iw_n), G0 from
    numpy.linalg.eigh of H(k) + P m0 P^H, accumulations and finalisation as in
    :264-383;
  GiwToNele (GreenFtn.cpp:513-635): the eigenvalue formula itself, with
    numpy.linalg.eigvals of H(k) + Sigma_n (intra-cluster blocks) per (k, n);
  two-pole iw -> tau transform: G(tau) = -1/2 sum_x e^{-x tau}/(1 + e^{-beta x});
  tau -> iw of a linear G(tau): G(iw_n) = c0 2i/w + c1 (i beta/w - 2/w^2).
The inputs are seeded (numpy default_rng); run from the repository root:
    python tests/golden/make_golden.py
"""
import os

import numpy as np

PI = 3.1415926535897932384626
HERE = os.path.dirname(os.path.abspath(__file__))


def herm(rng, d, scale=1.0):
    X = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    return scale * (X + X.conj().T) / 2


def lattice_case(seed, n, nc, ncl, nk, nf, beta, mu):
    rng = np.random.default_rng(seed)
    Hk = np.stack([herm(rng, n) for _ in range(nk)])
    # causal pole-form Sigma: sum_p A_p/(iw - b_p), A_p PSD
    w = PI * (2 * np.arange(nf) + 1) / beta
    sig = np.zeros((nf, nc, nc), complex)
    for _ in range(2):
        B = rng.normal(size=(nc, nc)) + 1j * rng.normal(size=(nc, nc))
        A = B @ B.conj().T / nc
        b = rng.uniform(-1, 1)
        sig += A[None] / (1j * w[:, None, None] - b)
    sig += herm(rng, nc, 0.3)[None]
    mom = np.stack([herm(rng, nc, s) for s in (0.3, 1.0, 0.5, 0.4)])
    return Hk, sig, mom


def sigma_at(sig, mom, n, nf, beta):
    if n < nf:
        return sig[n]
    z = 1j * PI * (2 * n + 1) / beta
    return mom[0] + mom[1] / z + mom[2] / z ** 2 + mom[3] / z ** 3


def greenftn_w_numpy(Hk, sig, mom, mu, beta, ncl, hf=2):
    nk, n, _ = Hk.shape
    nf, nc, _ = sig.shape
    npc = nc // ncl
    ev, U = np.linalg.eigh(Hk + np.pad(mom[0], ((0, n - nc), (0, n - nc)))[None])
    gloc = np.zeros((nf, nc, nc), complex)
    ghf = np.zeros((nf, nc, nc), complex)
    num = np.zeros((nc, nc), complex)
    dens = np.zeros((nk, n, n), complex)
    for nw in range(hf * nf):
        z = 1j * PI * (2 * nw + 1) / beta
        S = np.pad(sigma_at(sig, mom, nw, nf, beta), ((0, n - nc), (0, n - nc)))
        G = np.linalg.inv((z + mu) * np.eye(n)[None] - Hk - S[None])
        G0 = np.einsum("kil,kl,kjl->kij", U, 1.0 / (z + mu - ev), U.conj())
        if nw < nf:
            gloc[nw] = G[:, :nc, :nc].mean(0)
            ghf[nw] = G0[:, :nc, :nc].mean(0)
        D = G + np.conj(np.swapaxes(G, 1, 2)) - G0 - np.conj(np.swapaxes(G0, 1, 2))
        num += D[:, :nc, :nc].sum(0)
        dens += D
    x = beta * (ev - mu)
    f = np.where(x > 100, 0.0, np.where(x < -100, 1.0, 1.0 / (1.0 + np.exp(np.clip(x, -700, 700)))))
    occ = np.einsum("kil,kl,kjl->kij", U, f, U.conj())
    dens = dens / beta + occ
    num = (num / beta + occ[:, :nc, :nc].sum(0)) / nk
    mask = np.kron(np.eye(ncl), np.ones((npc, npc)))
    return gloc * mask, ghf * mask, num, dens


def giw_to_nele_numpy(Hk, sig, mom, mu, beta, ncl, hf=2):
    nk, n, _ = Hk.shape
    nf, nc, _ = sig.shape
    npc = nc // ncl
    mask = np.kron(np.eye(ncl), np.ones((npc, npc)))
    momm = mom * mask[None]
    linf = np.linalg.eigvals(Hk + np.pad(momm[0], ((0, n - nc), (0, n - nc)))[None])
    acc = 0.0
    for nw in range(hf * nf):
        wz = 1j * PI / beta + nw * (1j * 2 * PI / beta)
        S = sigma_at(sig * mask[None], momm, nw, nf, beta)
        lam = np.linalg.eigvals(Hk + np.pad(S, ((0, n - nc), (0, n - nc)))[None])
        acc += np.sum((1.0 / (wz - lam + mu) - 1.0 / (wz - linf + mu)).real)
    acc *= 2.0 / beta
    acc += np.sum(1.0 / (1.0 + np.exp(beta * (linf.real - mu))))
    return acc / nk


def main():
    cases = {
        "lat_n5_nc3": dict(seed=11, n=5, nc=3, ncl=1, nk=6, nf=8, beta=10.0, mu=0.2),
        "lat_n6_nc4_2cl": dict(seed=12, n=6, nc=4, ncl=2, nk=5, nf=10, beta=8.0, mu=-0.3),
        "lat_n3_nc3": dict(seed=13, n=3, nc=3, ncl=1, nk=7, nf=6, beta=12.0, mu=0.1),
        "lat_n8_nc2": dict(seed=14, n=8, nc=2, ncl=1, nk=4, nf=9, beta=6.0, mu=0.5),
    }
    for name, c in cases.items():
        Hk, sig, mom = lattice_case(c["seed"], c["n"], c["nc"], c["ncl"], c["nk"], c["nf"], c["beta"], c["mu"])
        gl, ghf, num, dens = greenftn_w_numpy(Hk, sig, mom, c["mu"], c["beta"], c["ncl"])
        N = giw_to_nele_numpy(Hk, sig, mom, c["mu"], c["beta"], c["ncl"])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), Hk=Hk, sigma=sig, moments=mom, mu=c["mu"],
                            beta=c["beta"], n_cluster=c["ncl"], gloc=gl, gloc_hf=ghf, num_matrix=num, density=dens,
                            n_total=N)
    # iw -> tau, two poles at a, b per diagonal entry (exact tail)
    beta, nf, ntau = 20.0, 64, 40
    a = np.array([-0.7, 0.2, 1.1])
    b = np.array([0.4, -1.3, 0.9])
    w = PI * (2 * np.arange(nf) + 1) / beta
    d = a.size
    fw = np.zeros((nf, d, d), complex)
    for i in range(d):
        fw[:, i, i] = 0.5 * (1 / (1j * w - a[i]) + 1 / (1j * w - b[i]))
    mom = np.zeros((4, d, d), complex)
    mom[1] = np.eye(d)
    mom[2] = np.diag((a + b) / 2)
    mom[3] = np.diag((a ** 2 + b ** 2) / 2)
    tau = np.arange(ntau + 1) * beta / ntau
    gt = np.zeros((ntau + 1, d, d), complex)
    for i in range(d):
        for x in (a[i], b[i]):
            gt[:, i, i] += -0.5 * np.exp(-x * tau) / (1 + np.exp(-beta * x))
    np.savez_compressed(os.path.join(HERE, "ft_two_pole.npz"), fw=fw, moments=mom, beta=beta, n_tau=ntau, gtau=gt)
    # tau -> iw of a linear function (exact for the piecewise-linear integral)
    c0 = np.array([[-0.3, 0.1 + 0.2j], [0.1 - 0.2j, -0.6]])
    c1 = np.array([[0.01, 0.0], [0.0, 0.02]])
    gt = c0[None] + c1[None] * tau[:, None, None]
    fw = 2j / w[:, None, None] * c0[None] + c1[None] * (1j * beta / w[:, None, None] - 2 / w[:, None, None] ** 2)
    np.savez_compressed(os.path.join(HERE, "ft_linear_tau.npz"), gtau=gt, fw=fw, beta=beta, n_freq=nf)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
