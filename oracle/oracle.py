"""ctypes front-end of the CPU oracle (oracle/dmft_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the dmftpack_amd package.

Arrays cross this boundary in "math" layout: a stack of matrices is a numpy
array of shape [count, rows, cols] with A[m, i, j] = A_m(i, j).  The C side is
column-major per matrix (Eigen MatrixXcd), so every call transposes the last
two axes on the way in and out.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libdmft_oracle.so")
_lib = None


class OrcLattice(C.Structure):
    _fields_ = [("n", C.c_int), ("n_corr", C.c_int), ("n_cluster", C.c_int), ("n_k", C.c_int),
                ("n_k_global", C.c_int), ("n_freq", C.c_int), ("high_freq", C.c_int), ("beta", C.c_double)]


class OrcCore(C.Structure):
    _fields_ = [("n_orb", C.c_int), ("eps", C.c_void_p), ("val_to_ks", C.c_void_p)]


def _core(core, nk, n):
    """core = (ks_energy [nk][n_orb] float, val_to_ks [nk][n] int, -1 = padded band) or None.
    Returns (ctypes pointer or None, arrays kept alive)."""
    if core is None:
        return None, None
    eps = np.ascontiguousarray(core[0], dtype=np.float64)
    v2k = np.ascontiguousarray(core[1], dtype=np.int32)
    assert eps.shape[0] == nk and v2k.shape == (nk, n)
    c = OrcCore(eps.shape[1], eps.ctypes.data_as(C.c_void_p), v2k.ctypes.data_as(C.c_void_p))
    return C.byref(c), (c, eps, v2k)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        _lib.orc_greenftn_w.argtypes = [C.POINTER(OrcLattice), P, P, P, C.c_double, P, P, P, P, C.c_int]
        _lib.orc_giw_to_nele.argtypes = [C.POINTER(OrcLattice), P, P, P, C.c_double, C.POINTER(OrcCore), C.c_int]
        _lib.orc_giw_to_nele.restype = C.c_double
        _lib.orc_adjust_mu.argtypes = [C.POINTER(OrcLattice), P, P, P, C.c_double, C.c_double, C.c_double,
                                       C.POINTER(OrcCore), C.c_int, C.POINTER(C.c_int)]
        _lib.orc_adjust_mu.restype = C.c_double
        _lib.orc_density.argtypes = [C.POINTER(OrcLattice), P, P, P, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(OrcCore), C.c_int]
        _lib.orc_core_bands.argtypes = [C.POINTER(OrcLattice), C.POINTER(OrcCore), C.c_double,
                                        C.POINTER(C.c_double)]
        _lib.orc_core_bands.restype = C.c_double
        _lib.orc_select_window.argtypes = [C.c_int, P, P]
        _lib.orc_causality_patch.argtypes = [C.c_int, C.c_int, C.c_double, P]
        _lib.orc_gloc_vjp.argtypes = [C.POINTER(OrcLattice), P, P, P, C.c_double, P, C.c_double, P,
                                      C.POINTER(C.c_double), C.c_int]
        _lib.orc_asymto_moments.argtypes = [C.c_int, C.c_int, C.c_double, P, C.c_int, P, P]
        _lib.orc_ft_tau.argtypes = [C.c_int, C.c_int, C.c_double, P, P, C.c_double, P]
        _lib.orc_ft_w2tau.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, P, P, P, C.POINTER(C.c_int), C.c_int]
        _lib.orc_ft_tau2w.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, P, P, C.POINTER(C.c_int)]
        _lib.orc_lu_inverse.argtypes = [C.c_int, P, P]
        _lib.orc_herm_eig.argtypes = [C.c_int, P, P, P]
        _lib.orc_num_threads.restype = C.c_int
        _lib.orc_hyb_delta.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, P, P, P, C.c_int, P, P]
        _lib.orc_hk_from_hr.argtypes = [C.c_int, C.c_int, P, P, C.c_int, P, P]
        _lib.orc_greenftn_real.argtypes = [C.POINTER(OrcLattice), P, C.c_int, C.c_double, C.c_double, C.c_double,
                                           P, C.c_double, P, P, C.c_int]
        _lib.orc_upfold_density.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P, P,
                                            P, P, P, P, C.c_int, P, P, P, P]
    return _lib


def _cm(a):
    """math layout -> contiguous column-major complex128 buffer"""
    a = np.asarray(a, dtype=np.complex128)
    return np.ascontiguousarray(np.swapaxes(a, -1, -2))


def _from_cm(buf):
    return np.ascontiguousarray(np.swapaxes(buf, -1, -2))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _lat(n, n_corr, n_cluster, n_k, n_k_global, n_freq, beta, high_freq=2):
    return OrcLattice(n, n_corr, n_cluster, n_k, n_k_global, n_freq, high_freq, beta)


def num_threads():
    return lib().orc_num_threads()


def lu_inverse(A):
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    a = _cm(A)
    out = np.empty_like(a)
    lib().orc_lu_inverse(n, _p(a), _p(out))
    return _from_cm(out)


def herm_eig(A):
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    a = _cm(A)
    w = np.empty(n)
    v = np.empty_like(a)
    lib().orc_herm_eig(n, _p(a), _p(w), _p(v))
    return w, _from_cm(v)


def greenftn_w(Hk, sigma, moments, mu, beta, n_cluster=1, n_k_global=None, density=False, nthreads=1, high_freq=2):
    """GreenFtn_w restated; returns dict(gloc, gloc_hf, num_matrix[, density])."""
    Hk = np.asarray(Hk, dtype=np.complex128)
    nk, n, _ = Hk.shape
    nf, nc, _ = sigma.shape
    L = _lat(n, nc, n_cluster, nk, n_k_global or nk, nf, beta, high_freq)
    h, s, m = _cm(Hk), _cm(sigma), _cm(moments)
    gl = np.zeros((nf, nc, nc), np.complex128)
    gh = np.zeros_like(gl)
    nm = np.zeros((nc, nc), np.complex128)
    dm = np.zeros((nk, n, n), np.complex128) if density else None
    lib().orc_greenftn_w(C.byref(L), _p(h), _p(s), _p(m), mu, _p(gl), _p(gh), _p(nm), _p(dm), nthreads)
    out = dict(gloc=_from_cm(gl), gloc_hf=_from_cm(gh), num_matrix=_from_cm(nm))
    if density:
        out["density"] = _from_cm(dm)
    return out


def giw_to_nele(Hk, sigma, moments, mu, beta, n_cluster=1, nthreads=1, high_freq=2, core=None):
    """GiwToNele restated; core = (KS_eigenEnergy [nk][NumOrbit], FromValToKS [nk][n]) adds the
    core-band term (GreenFtn.cpp:622-631)."""
    Hk = np.asarray(Hk, dtype=np.complex128)
    nk, n, _ = Hk.shape
    nf, nc, _ = sigma.shape
    L = _lat(n, nc, n_cluster, nk, nk, nf, beta, high_freq)
    h, s, m = _cm(Hk), _cm(sigma), _cm(moments)
    cp, _keep = _core(core, nk, n)
    return lib().orc_giw_to_nele(C.byref(L), _p(h), _p(s), _p(m), mu, cp, nthreads)


def adjust_mu(Hk, sigma, moments, mu, beta, n_target, dmu=1.0, n_cluster=1, nthreads=1, core=None):
    Hk = np.asarray(Hk, dtype=np.complex128)
    nk, n, _ = Hk.shape
    nf, nc, _ = sigma.shape
    L = _lat(n, nc, n_cluster, nk, nk, nf, beta)
    h, s, m = _cm(Hk), _cm(sigma), _cm(moments)
    cp, _keep = _core(core, nk, n)
    ne = C.c_int(0)
    mu_out = lib().orc_adjust_mu(C.byref(L), _p(h), _p(s), _p(m), mu, dmu, n_target, cp, nthreads, C.byref(ne))
    return mu_out, ne.value


def density(Hk, sigma, moments, mu, beta, nthreads=1, high_freq=2, core=None):
    Hk = np.asarray(Hk, dtype=np.complex128)
    nk, n, _ = Hk.shape
    nf, nc, _ = sigma.shape
    L = _lat(n, nc, 1, nk, nk, nf, beta, high_freq)
    h, s, m = _cm(Hk), _cm(sigma), _cm(moments)
    cp, _keep = _core(core, nk, n)
    N, dN = C.c_double(0), C.c_double(0)
    lib().orc_density(C.byref(L), _p(h), _p(s), _p(m), mu, C.byref(N), C.byref(dN), cp, nthreads)
    return N.value, dN.value


def core_bands(ks_energy, val_to_ks, mu, beta, n_k_global=None):
    """Core-band term of GiwToNele alone, normalised by n_k_global; returns (term, d term/dmu)."""
    eps = np.asarray(ks_energy, dtype=np.float64)
    v2k = np.asarray(val_to_ks, dtype=np.int32)
    nk, n = v2k.shape
    L = _lat(n, 1, 1, nk, n_k_global or nk, 1, beta)
    cp, _keep = _core((eps, v2k), nk, n)
    d = C.c_double(0)
    s = lib().orc_core_bands(C.byref(L), cp, mu, C.byref(d))
    return s / L.n_k_global, d.value / L.n_k_global


def gloc_vjp(Hk, sigma, moments, mu, beta, gbar, nbar, nthreads=1):
    Hk = np.asarray(Hk, dtype=np.complex128)
    nk, n, _ = Hk.shape
    nf, nc, _ = sigma.shape
    L = _lat(n, nc, 1, nk, nk, nf, beta)
    h, s, m, g = _cm(Hk), _cm(sigma), _cm(moments), _cm(gbar)
    sb = np.zeros_like(s)
    mb = C.c_double(0)
    lib().orc_gloc_vjp(C.byref(L), _p(h), _p(s), _p(m), mu, _p(g), nbar, _p(sb), C.byref(mb), nthreads)
    return _from_cm(sb), mb.value


def asymto_moments(fw, beta, green_asym=False, return_windows=False):
    fw = np.asarray(fw, dtype=np.complex128)
    nf, d, _ = fw.shape
    f = _cm(fw)
    mom = np.zeros((4, d, d), np.complex128)
    win = (C.c_int * 4)()
    rc = lib().orc_asymto_moments(d, nf, beta, _p(f), int(green_asym), _p(mom), win)
    if rc != 0:
        raise ValueError("getAsymto_moments: too few frequencies for the window search")
    m = _from_cm(mom)
    return (m, list(win)) if return_windows else m


def ft_tau(fw, moments, beta, tau):
    fw = np.asarray(fw, dtype=np.complex128)
    nf, d, _ = fw.shape
    f, m = _cm(fw), _cm(moments)
    out = np.zeros((d, d), np.complex128)
    lib().orc_ft_tau(d, nf, beta, _p(f), _p(m), tau, _p(out))
    return _from_cm(out)


def ft_w2tau(fw, moments, beta, n_tau, nthreads=1):
    fw = np.asarray(fw, dtype=np.complex128)
    nf, d, _ = fw.shape
    f, m = _cm(fw), _cm(moments)
    out = np.zeros((n_tau + 1, d, d), np.complex128)
    chk = C.c_int(0)
    lib().orc_ft_w2tau(d, nf, n_tau, beta, _p(f), _p(m), _p(out), C.byref(chk), nthreads)
    return _from_cm(out), chk.value


def ft_tau2w(gt, beta, n_freq):
    gt = np.asarray(gt, dtype=np.complex128)
    nt1, d, _ = gt.shape
    g = _cm(gt)
    out = np.zeros((n_freq, d, d), np.complex128)
    viol = C.c_int(0)
    lib().orc_ft_tau2w(d, n_freq, nt1 - 1, beta, _p(g), _p(out), C.byref(viol))
    return _from_cm(out), viol.value


def hyb_delta(gloc, sigma, h_imp, orbit, beta, mu):
    """Construct_hyb_delta restated: Delta(iw_n) of the impurity `orbit` (indices into the correlated block)."""
    gloc = np.asarray(gloc, dtype=np.complex128)
    nf, nc, _ = gloc.shape
    orbit = np.ascontiguousarray(np.asarray(orbit, dtype=np.int32))
    dim = orbit.size
    out = np.zeros((nf, dim, dim), np.complex128)
    g, sg, hi = _cm(gloc), _cm(sigma), _cm(np.asarray(h_imp)[None])
    rc = lib().orc_hyb_delta(nc, nf, beta, mu, _p(g), _p(sg), _p(hi), dim, _p(orbit), _p(out))
    if rc:
        raise ValueError("singular G_imp")
    return _from_cm(out)


def hk_from_hr(n, hops, kphase):
    """H(k) = sum_R H(R) exp(-i k.a.R) from a hopping list [(R1, R2, R3, i, j, t), ...]; kphase [nk][3]."""
    idx = np.ascontiguousarray(np.array([h[:5] for h in hops], dtype=np.int32).reshape(-1, 5))
    t = np.ascontiguousarray(np.array([complex(h[5]) for h in hops], dtype=np.complex128))
    kp = np.ascontiguousarray(np.asarray(kphase, dtype=np.float64))
    nk = kp.shape[0]
    out = np.zeros((nk, n, n), np.complex128)
    lib().orc_hk_from_hr(n, len(hops), _p(idx), _p(t), nk, _p(kp), _p(out))
    return _from_cm(out)


def greenftn_real(Hk, sigma_r, mu, e0, de, eta, nthreads=1):
    """retarded_GreenFtn2 summed over k: (G^R_loc [n_e][nc][nc], (1/Nk) sum_k tr G^R [n_e])."""
    Hk = np.asarray(Hk, dtype=np.complex128)
    sigma_r = np.asarray(sigma_r, dtype=np.complex128)
    nk, n, _ = Hk.shape
    ne, nc, _ = sigma_r.shape
    L = _lat(n, nc, 1, nk, nk, ne, 1.0, 1)
    h, s = _cm(Hk), _cm(sigma_r)
    g = np.zeros((ne, nc, nc), np.complex128)
    tr = np.zeros(ne, np.complex128)
    rc = lib().orc_greenftn_real(C.byref(L), _p(h), ne, e0, de, eta, _p(s), mu, _p(g), _p(tr), nthreads)
    if rc:
        raise ValueError("singular matrix")
    return _from_cm(g), tr


def upfold_density(dm_model, corr_base, ks_energy, val_to_ks, transform, ks_ortho, overlap, beta, mu, r_idx, kphase,
                   n_k_global=None):
    """upfolding_density with upfolding_k restated (downfolding == 1):
    returns (D [nk][n_orb][n_orb], dmr [3][n_r]) with dmr = (dual-dual, dual-direct, direct-direct) DM(R).
    overlap None means S = I."""
    X = np.asarray(dm_model, dtype=np.complex128)
    nk, n, _ = X.shape
    eps = np.ascontiguousarray(np.asarray(ks_energy, dtype=np.float64))
    n_orb = eps.shape[1]
    v = np.ascontiguousarray(np.asarray(val_to_ks, dtype=np.int32))
    ri = np.ascontiguousarray(np.asarray(r_idx, dtype=np.int32).reshape(-1, 5))
    kp = np.ascontiguousarray(np.asarray(kphase, dtype=np.float64))
    x, c, t, ko = _cm(X), _cm(corr_base), _cm(transform), _cm(ks_ortho)
    s = None if overlap is None else _cm(overlap)
    dm = np.zeros((nk, n_orb, n_orb), np.complex128)
    dmr = np.zeros((3, ri.shape[0]), np.complex128)
    rc = lib().orc_upfold_density(n, n_orb, nk, n_k_global or nk, beta, mu, _p(x), _p(c), _p(eps), _p(v), _p(t),
                                  _p(ko), _p(s), ri.shape[0], _p(ri), _p(kp), _p(dm), _p(dmr))
    if rc:
        raise ValueError("singular transform or overlap")
    return _from_cm(dm), dmr


def select_window(var, ok):
    """The window scan of getAsymto_moments on given spreads and positivity flags: returns minPos."""
    v = np.ascontiguousarray(var, dtype=np.float64)
    o = np.ascontiguousarray(ok, dtype=np.int32)
    return lib().orc_select_window(v.size, _p(v), _p(o))


def causality_patch(gt, beta):
    """FourierTransform_'s causality patch on gt [n_tau+1][d][d]: returns (patched copy, CHECK)."""
    g = _cm(gt)
    check = lib().orc_causality_patch(g.shape[1], g.shape[0] - 1, beta, _p(g))
    return _from_cm(g), check
