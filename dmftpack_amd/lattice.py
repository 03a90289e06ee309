"""Synthetic lattice inputs: k-mesh, H(k) models and pole-form self-energies.

Everything here is deterministic from a seed through a counter-based PRNG
(splitmix64), so the CPU oracle and the HIP path see bit-identical inputs.

Conventions follow the reference:
  * k-mesh (TB.cpp:196-206): k = kx + ky*Nx + kz*Nx*Ny,
    k_i = -pi/a + delta/2 + j*delta, delta = 2 pi/(a N_i), a = ax = pi
    (dmft_common.h:37), so the phase k_i*a = -pi + pi/N_i + 2 pi j/N_i.
  * H(k) = sum_R H(R) exp(-i k.a R)   (ConstructModelHamiltonian.cpp:58-77).
  * Matsubara frequencies w_n = pi (2n+1)/beta (GreenFtn.cpp:219).
  * Correlated orbitals are the first n_corr of the basis (downfolding.cpp:69-79).
Arrays are in "math" layout: H[k, i, j] = H_k(i, j).
"""
import numpy as np

PI = 3.1415926535897932384626  # dmft_common.h:43

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, count):
    """count doubles in [0, 1) from (seed, stream) -- counter-based splitmix64."""
    with np.errstate(over="ignore"):
        base = _mix(np.array([seed], dtype=np.uint64) * _GOLD + np.uint64(stream) * np.uint64(0xD1B54A32D192ED03))
        ctr = base + (np.arange(1, count + 1, dtype=np.uint64) * _GOLD)
        x = _mix(ctr)
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, count):
    """count standard normals (Box-Muller on two uniform streams)."""
    u1 = uniform(seed, 2 * stream + 1, count)
    u2 = uniform(seed, 2 * stream + 2, count)
    u1 = np.where(u1 <= 0.0, 1e-300, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * PI * u2)


def kmesh_phases(nx, ny=None, nz=None):
    """[Nk, 3] phases k_i * a on the reference's shifted uniform mesh."""
    ny = ny or nx
    nz = nz or nx
    k = np.arange(nx * ny * nz)
    jx, jy, jz = k % nx, (k % (nx * ny)) // nx, k // (nx * ny)
    th = np.empty((k.size, 3))
    for c, (j, m) in enumerate(((jx, nx), (jy, ny), (jz, nz))):
        delta = 2.0 * PI / (PI * m)
        th[:, c] = (-PI / PI + delta / 2.0 + j * delta) * PI
    return th


def matsubara(n_freq, beta):
    return PI * (2.0 * np.arange(n_freq) + 1.0) / beta


# --------------------------------------------------------------------------
# H(k) models
# --------------------------------------------------------------------------
def t2g_hk(nx, ny=None, nz=None, t=0.25, tz=0.03):
    """3-orbital t2g cubic model, orbital-diagonal H(k)."""
    ny = ny or nx
    nz = nz or nx
    th = kmesh_phases(nx, ny, nz)
    cx, cy, cz = np.cos(th[:, 0]), np.cos(th[:, 1]), np.cos(th[:, 2])
    eps = np.stack([-2 * t * (cx + cy) - tz * cz,
                    -2 * t * (cy + cz) - tz * cx,
                    -2 * t * (cx + cz) - tz * cy], axis=1)
    Hk = np.zeros((th.shape[0], 3, 3), np.complex128)
    for a in range(3):
        Hk[:, a, a] = eps[:, a]
    return Hk, eps


# one representative of each +-R pair in the first two cubic shells
_SHELL_HALF = [(1, 0, 0), (0, 1, 0), (0, 0, 1),
               (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1)]


def random_tb_hr(n, seed, sigma_t=0.5, onsite=(-3.0, 2.0)):
    """Seeded Hermitian tight binding: returns (R list, H(R) list) incl. R=0."""
    Rs, HR = [], []
    eps = onsite[0] + (onsite[1] - onsite[0]) * uniform(seed, 1, n)
    g = normal(seed, 2, 2 * n * n).reshape(2, n, n)
    X = (g[0] + 1j * g[1]) * (sigma_t / np.sqrt(2.0))
    H0 = np.triu(X, 1)
    H0 = H0 + H0.conj().T + np.diag(eps)
    Rs.append((0, 0, 0))
    HR.append(H0)
    for r, R in enumerate(_SHELL_HALF):
        var = sigma_t ** 2 * np.exp(-np.linalg.norm(R))
        g = normal(seed, 10 + r, 2 * n * n).reshape(2, n, n)
        HR.append((g[0] + 1j * g[1]) * np.sqrt(var / 2.0))
        Rs.append(R)
    return Rs, HR


def hk_from_hr(Rs, HR, nx, ny=None, nz=None):
    """H(k) = H(0) + sum_{R in half shells} [H(R) e^{-ik.R} + H(R)^H e^{+ik.R}]."""
    ny = ny or nx
    nz = nz or nx
    th = kmesh_phases(nx, ny, nz)
    n = HR[0].shape[0]
    Hk = np.broadcast_to(HR[0], (th.shape[0], n, n)).astype(np.complex128).copy()
    for R, H in zip(Rs[1:], HR[1:]):
        ph = np.exp(-1j * (th @ np.asarray(R, dtype=np.float64)))
        Hk += ph[:, None, None] * H[None] + ph.conj()[:, None, None] * H.conj().T[None]
    return Hk


def random_tb_hk(n, nx, seed, ny=None, nz=None):
    Rs, HR = random_tb_hr(n, seed)
    return hk_from_hr(Rs, HR, nx, ny, nz)


def hopping_list(Rs, HR, drop_zero=True):
    """Full hopping list [(R1, R2, R3, i, j, t), ...] (0-based orbitals) of a
    half-shell (Rs, HR) model: each R != 0 also contributes H(-R) = H(R)^H, the
    form an Hk.HWR file lists (read_TB_parm.cpp:36-62)."""
    hops = []
    for R, H in zip(Rs, HR):
        n = H.shape[0]
        parts = [(R, H)] if tuple(R) == (0, 0, 0) else [(R, H), (tuple(-np.asarray(R)), H.conj().T)]
        for RR, HH in parts:
            for j in range(n):
                for i in range(n):
                    if drop_zero and HH[i, j] == 0:
                        continue
                    hops.append((int(RR[0]), int(RR[1]), int(RR[2]), i, j, complex(HH[i, j])))
    return hops


def read_hwr(path):
    """Hk.HWR text reader (read_HR, read_TB_parm.cpp:6-76): the header lines
    "Norbital of Atom = N" (one per atom), then one line per hopping
    "R1 R2 R3 atom1 atom2 i j Re Im" with 1-based atoms and orbitals.
    Returns (orbitals per atom, hops [(R1, R2, R3, i0, j0, t)]) with
    i0 = (orbitals of the atoms before atom1) + i - 1, likewise j0."""
    norbs, hops = [], []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s:
                continue
            if s.startswith("Norbital of Atom"):
                norbs.append(int(s.split("=")[1]))
                continue
            f = s.split()
            R = [int(v) for v in f[:3]]
            a1, a2, i, j = (int(v) for v in f[3:7])
            acc = np.concatenate([[0], np.cumsum(norbs)])
            hops.append((R[0], R[1], R[2], int(acc[a1 - 1]) + i - 1, int(acc[a2 - 1]) + j - 1,
                         float(f[7]) + 1j * float(f[8])))
    return norbs, hops


def t2g_hopping_list(t=0.25, tz=0.03):
    """Hopping list of the t2g model of t2g_hk (orbital xy: -t along x and y,
    -tz/2 along z; yz and xz likewise)."""
    axes = {0: ((1, 0, 0), (0, 1, 0), (0, 0, 1)), 1: ((0, 1, 0), (0, 0, 1), (1, 0, 0)),
            2: ((1, 0, 0), (0, 0, 1), (0, 1, 0))}
    hops = []
    for a, (u, v, w) in axes.items():
        for R, amp in ((u, -t), (v, -t), (w, -tz / 2.0)):
            for sgn in (1, -1):
                hops.append((sgn * R[0], sgn * R[1], sgn * R[2], a, a, complex(amp)))
    return hops


# --------------------------------------------------------------------------
# Self-energies
# --------------------------------------------------------------------------
def t2g_sigma(n_freq, beta, seed, n_orb=3):
    """Diagonal Sigma_a(iw) = S_inf + A/(iw + i Gamma), S_inf~U[0,2], A~U[.5,2], Gamma~U[.2,1]."""
    u = uniform(seed, 100, 3 * n_orb).reshape(3, n_orb)
    sinf, A, gam = 2.0 * u[0], 0.5 + 1.5 * u[1], 0.2 + 0.8 * u[2]
    w = matsubara(n_freq, beta)
    sig = np.zeros((n_freq, n_orb, n_orb), np.complex128)
    for a in range(n_orb):
        sig[:, a, a] = sinf[a] + A[a] / (1j * w + 1j * gam[a])
    return sig, dict(sinf=sinf, A=A, gamma=gam)


def pole_sigma(n_corr, n_freq, beta, seed, n_poles=3, block=None):
    """Sigma(iw) = sum_p A_p/(iw - b_p), A_p = B B^H/5 (PSD => causal), b_p~U[-2,2].

    block: if given, Sigma is block-diagonal with independent blocks of that
    size.  Returns (sigma[n_freq,nc,nc], exact tail moments[4,nc,nc])."""
    block = block or n_corr
    w = matsubara(n_freq, beta)
    sig = np.zeros((n_freq, n_corr, n_corr), np.complex128)
    mom = np.zeros((4, n_corr, n_corr), np.complex128)
    for bi, s0 in enumerate(range(0, n_corr, block)):
        d = min(block, n_corr - s0)
        for p in range(n_poles):
            g = normal(seed, 200 + 10 * bi + p, 2 * d * d).reshape(2, d, d)
            B = (g[0] + 1j * g[1]) / np.sqrt(2.0)
            A = B @ B.conj().T / 5.0
            b = -2.0 + 4.0 * uniform(seed, 300 + 10 * bi + p, 1)[0]
            sl = slice(s0, s0 + d)
            sig[:, sl, sl] += A[None] / (1j * w[:, None, None] - b)
            mom[1, sl, sl] += A
            mom[2, sl, sl] += A * b
            mom[3, sl, sl] += A * b * b
    return sig, mom


def perturbed(sig, seed, scale=0.05):
    """Seeded multiplicative perturbation that keeps causality (used for G_tgt)."""
    nf, d, _ = sig.shape
    u = uniform(seed, 400, d)
    D = np.diag(1.0 + scale * (2 * u - 1))
    return np.einsum("ij,njk,kl->nil", D, sig, D)


# --------------------------------------------------------------------------
# Named workloads (sizes of the headline configurations)
# --------------------------------------------------------------------------
def workload(name, seed=1, nk_side=None, n_freq=None):
    """Build one named workload.  nk_side / n_freq override the sizes (tests)."""
    if name == "t2g":  # 3-orbital t2g, Nk=16^3, beta=40, Nw=1024
        nx = nk_side or 16
        nf = n_freq or 1024
        beta = 40.0
        Hk, _ = t2g_hk(nx)
        sig, _ = t2g_sigma(nf, beta, seed)
        return dict(name=name, Hk=Hk, sigma=sig, beta=beta, n_cluster=1, n_target=0.5, mu=0.0,
                    n_tau=2048, nk_side=nx)
    if name == "dp14":  # 14-orbital d+p, 5x5 correlated block, Nk=32^3, Nw=2048, beta=50
        nx = nk_side or 32
        nf = n_freq or 2048
        beta = 50.0
        Hk = random_tb_hk(14, nx, seed)
        sig, mom = pole_sigma(5, nf, beta, seed)
        return dict(name=name, Hk=Hk, sigma=sig, exact_moments=mom, beta=beta, n_cluster=1, n_target=7.0,
                    mu=0.0, n_tau=2048, nk_side=nx)
    if name == "dp14_grad":  # the 14-orbital system at Nk=16^3, Nw=1024, beta=40
        nx = nk_side or 16
        nf = n_freq or 1024
        beta = 40.0
        Hk = random_tb_hk(14, nx, seed)
        sig, mom = pole_sigma(5, nf, beta, seed)
        return dict(name=name, Hk=Hk, sigma=sig, exact_moments=mom, beta=beta, n_cluster=1, n_target=7.0,
                    mu=0.0, n_tau=2048, nk_side=nx)
    if name == "large32":  # 32 orbitals, 2 correlated atoms x 5 d, Nk=40^3, Nw=4096, beta=100
        nx = nk_side or 40
        nf = n_freq or 4096
        beta = 100.0
        Hk = random_tb_hk(32, nx, seed)
        sig, mom = pole_sigma(10, nf, beta, seed, block=5)
        return dict(name=name, Hk=Hk, sigma=sig, exact_moments=mom, beta=beta, n_cluster=2, n_target=16.0,
                    mu=0.0, n_tau=8192, nk_side=nx)
    raise KeyError(name)


def batch_systems(count=64, seed=1, n_freq=1024, beta=40.0, norbs=(3, 5, 7, 10, 14, 20), nk_sides=(8, 12, 16, 24)):
    """Independent systems of uneven size for the batched path."""
    systems = []
    for s in range(count):
        u = uniform(seed + s, 500, 2)
        n = norbs[int(u[0] * len(norbs))]
        nx = nk_sides[int(u[1] * len(nk_sides))]
        nc = min(5, n)
        Hk = random_tb_hk(n, nx, seed + s)
        sig, mom = pole_sigma(nc, n_freq, beta, seed + s)
        systems.append(dict(Hk=Hk, sigma=sig, exact_moments=mom, beta=beta, n_cluster=1, nk_side=nx, n=n, n_corr=nc,
                            mu=0.0))
    return systems


def _cnormal(seed, stream, shape):
    g = normal(seed, stream, 2 * int(np.prod(shape))).reshape((2,) + tuple(shape))
    return (g[0] + 1j * g[1]) / np.sqrt(2.0)


def _unitary(seed, stream, count, d):
    q, r = np.linalg.qr(_cnormal(seed, stream, (count, d, d)))
    ph = np.diagonal(r, axis1=1, axis2=2)
    return q * (ph / np.abs(ph))[:, None, :]


def upfold_case(n, n_orb, nk_side, seed, overlap=True, identity=False):
    """Synthetic inputs of upfolding_density (upfolding.cpp:38-86):
    dm_model [nk][n][n] Hermitian, corr_base [nk][n][n] unitary (DF_CorrBase),
    ks_energy [nk][n_orb], val_to_ks [nk][n] distinct band indices (FromValToKS),
    transform [nk][n_orb][n_orb] (transformMatrix_k, well conditioned),
    ks_ortho [nk][n_orb][n_orb] unitary (KS_eigenVectors_orthoBasis),
    overlap [nk][n_orb][n_orb] Hermitian positive definite or None,
    r_idx [n_r][5] = every (R, i0, m0) with R in {-1,0,1}^3, kphase [nk][3].
    identity=True sets every basis change to the identity (n = n_orb)."""
    kp = kmesh_phases(nk_side)
    nk = kp.shape[0]
    A = _cnormal(seed, 501, (nk, n, n))
    X = 0.5 * (A + A.conj().transpose(0, 2, 1)) / np.sqrt(n)
    eps = np.sort(-4.0 + 8.0 * uniform(seed, 502, nk * n_orb).reshape(nk, n_orb), axis=1)
    eps[:, 0] = -30.0  # beta (eps - mu) beyond -100 and +100: the clamped Fermi factors
    eps[:, -1] = 30.0
    if identity:
        assert n == n_orb
        eye_n = np.broadcast_to(np.eye(n, dtype=complex), (nk, n, n)).copy()
        eye_o = eye_n.copy()
        return dict(dm_model=X, corr_base=eye_n, ks_energy=eps, val_to_ks=np.tile(np.arange(n), (nk, 1)),
                    transform=eye_o, ks_ortho=eye_o.copy(), overlap=None if not overlap else eye_o.copy(),
                    r_idx=_all_r_idx(n_orb), kphase=kp)
    C = _unitary(seed, 503, nk, n)
    order = np.argsort(uniform(seed, 504, nk * n_orb).reshape(nk, n_orb), axis=1)
    v = np.ascontiguousarray(order[:, :n]).astype(np.int32)
    T = np.eye(n_orb)[None] + 0.3 * _cnormal(seed, 505, (nk, n_orb, n_orb)) / np.sqrt(n_orb)
    Ko = _unitary(seed, 506, nk, n_orb)
    S = None
    if overlap:
        B = _cnormal(seed, 507, (nk, n_orb, n_orb))
        S = np.eye(n_orb)[None] + 0.3 * (B @ B.conj().transpose(0, 2, 1)) / n_orb
    return dict(dm_model=X, corr_base=C, ks_energy=eps, val_to_ks=v, transform=T, ks_ortho=Ko, overlap=S,
                r_idx=_all_r_idx(n_orb), kphase=kp)


def _all_r_idx(n_orb):
    rows = []
    for r1 in (-1, 0, 1):
        for r2 in (-1, 0, 1):
            for r3 in (-1, 0, 1):
                for i in range(n_orb):
                    for j in range(n_orb):
                        rows.append((r1, r2, r3, i, j))
    return np.array(rows, dtype=np.int32)


def windowed_case(n, n_orb, nk_side, seed, n_below=2, padded=0):
    """A downfolded lattice with Kohn-Sham bands outside the model window
    (downfolding.cpp:162-175, GiwToNele's core-band term GreenFtn.cpp:622-631):
    ks_energy [nk][n_orb] ascending per k, the window being bands
    n_below .. n_below+n-1 (val_to_ks [nk][n], the window band order shuffled
    per k); H(k) = U diag(window energies) U^H with a seeded unitary U, so
    eig H(k) are the window's KS energies.  The bands below the window lie in
    [-1.6, -0.4], those above in [0.6, 1.8], close enough to mu ~ 0 that their
    Fermi factors and derivatives matter at beta ~ 10.  padded > 0 marks the
    last `padded` window bands of every other k as absent (val_to_ks = -1, the
    shim's padding of a smaller NBAND[k]): their H(k) rows and columns become
    decoupled levels at 1e4."""
    kp = kmesh_phases(nk_side)
    nk = kp.shape[0]
    n_above = n_orb - n - n_below
    assert n_above >= 0
    u = uniform(seed, 601, nk * n_orb).reshape(nk, n_orb)
    eps = np.empty((nk, n_orb))
    eps[:, :n_below] = -1.6 + 1.2 * u[:, :n_below]
    eps[:, n_below:n_below + n] = -0.8 + 1.6 * u[:, n_below:n_below + n]
    eps[:, n_below + n:] = 0.6 + 1.2 * u[:, n_below + n:]
    eps = np.sort(eps, axis=1)
    perm = np.argsort(uniform(seed, 602, nk * n).reshape(nk, n), axis=1)
    v2k = (n_below + perm).astype(np.int32)
    U = _unitary(seed, 603, nk, n)
    win = np.take_along_axis(eps, v2k, axis=1)
    if padded:
        for k in range(1, nk, 2):
            v2k[k, n - padded:] = -1
            win[k, n - padded:] = 1.0e4
            U[k, :, n - padded:] = 0.0
            U[k, n - padded:, :] = 0.0
            U[k, n - padded:, n - padded:] = np.eye(padded)
            q, r = np.linalg.qr(U[k, :n - padded, :n - padded])
            U[k, :n - padded, :n - padded] = q
    Hk = (U * win[:, None, :]) @ U.conj().transpose(0, 2, 1)
    return dict(Hk=0.5 * (Hk + Hk.conj().transpose(0, 2, 1)), ks_energy=eps, val_to_ks=v2k, kphase=kp)
