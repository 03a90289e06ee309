"""Thin torch-memory wrapper over the C ABI (include/dmftgpu.h).

Torch only provides device memory and the stream here; every computation runs
in libdmftgpu.so.  Arrays are accepted in "math" layout (A[m, i, j] = A_m(i, j))
as numpy arrays or torch tensors and converted to the column-major complex128
device buffers the ABI expects.  Method names follow the reference routines
(GreenFtn_w, GreenFtn_w_adjust_mu, getAsymto_moments, FourierTransform,
FT_t_to_w).
"""
import ctypes as C

import numpy as np
import torch

from . import capi

PLAIN_F64, SCHUR_F64, SCHUR_F32R = capi.PLAIN_F64, capi.SCHUR_F64, capi.SCHUR_F32R


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("dmftpack_amd needs a ROCm GPU: torch.cuda.is_available() is False")


def to_dev_cm(a, device="cuda"):
    """math-layout array -> contiguous column-major complex128 device tensor"""
    if isinstance(a, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(np.swapaxes(np.asarray(a, np.complex128), -1, -2)))
        return t.to(device)
    t = a.to(device=device, dtype=torch.complex128)
    return t.transpose(-1, -2).contiguous()


def from_dev_cm(t):
    """column-major device tensor -> math-layout numpy array"""
    return np.ascontiguousarray(np.swapaxes(t.cpu().numpy(), -1, -2))


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream_ptr(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Handle:
    def __init__(self, stream=None):
        _require_gpu()
        self.lib = capi.lib()
        self.h = C.c_void_p()
        self.stream = stream
        rc = self.lib.dmftgpu_create(C.byref(self.h), _stream_ptr(stream))
        if rc != 0:
            raise capi.DmftGpuError(rc, "dmftgpu_create failed")

    def check(self, rc):
        return capi.check(self.h, rc)

    def set_option(self, option, value):
        self.check(self.lib.dmftgpu_set_option(self.h, option, value))

    def enable_timing(self, on=True):
        self.check(self.lib.dmftgpu_enable_timing(self.h, int(on)))

    def kernel_ms(self, mode):
        ms = C.c_double(0)
        self.check(self.lib.dmftgpu_kernel_time(self.h, mode, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            self.lib.dmftgpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- Matsubara tails and transforms ----------------------------------
    def get_asymto_moments(self, fw, beta, green_asym=False, return_windows=False):
        """getAsymto_moments (FourierTransform.cpp:184-465)."""
        f = to_dev_cm(fw)
        nf, d = f.shape[0], f.shape[1]
        mom = torch.zeros((4, d, d), dtype=torch.complex128, device=f.device)
        win = (C.c_int * 4)()
        self.check(self.lib.dmftgpu_moments(self.h, d, nf, beta, _ptr(f), int(green_asym), _ptr(mom), win))
        m = from_dev_cm(mom)
        return (m, list(win)) if return_windows else m

    def fourier_transform(self, fw, moments, beta, n_tau):
        """FourierTransform_ (FourierTransform.cpp:126-178): G(iw) -> G(tau_j), j=0..n_tau."""
        f, m = to_dev_cm(fw), to_dev_cm(moments)
        nf, d = f.shape[0], f.shape[1]
        out = torch.zeros((n_tau + 1, d, d), dtype=torch.complex128, device=f.device)
        patched = C.c_int(0)
        self.check(self.lib.dmftgpu_ft_w2tau(self.h, d, nf, n_tau, beta, _ptr(f), _ptr(m), _ptr(out),
                                             C.byref(patched)))
        return from_dev_cm(out), patched.value

    def fourier_transform_tau(self, fw, moments, beta, taus):
        """FourierTransform_tau (FourierTransform.cpp:16-118) at the given tau values."""
        f, m = to_dev_cm(fw), to_dev_cm(moments)
        nf, d = f.shape[0], f.shape[1]
        taus = np.ascontiguousarray(np.asarray(taus, dtype=np.float64))
        out = torch.zeros((taus.size, d, d), dtype=torch.complex128, device=f.device)
        self.check(self.lib.dmftgpu_ft_w2tau_points(self.h, d, nf, beta, _ptr(f), _ptr(m), taus.size,
                                                    taus.ctypes.data_as(C.POINTER(C.c_double)), _ptr(out)))
        return from_dev_cm(out)

    def ft_t_to_w(self, gt, beta, n_freq):
        """FT_t_to_w (FourierTransform.cpp:521-561)."""
        g = to_dev_cm(gt)
        d = g.shape[1]
        out = torch.zeros((n_freq, d, d), dtype=torch.complex128, device=g.device)
        viol = C.c_int(0)
        self.check(self.lib.dmftgpu_ft_tau2w(self.h, d, n_freq, g.shape[0] - 1, beta, _ptr(g), _ptr(out),
                                             C.byref(viol)))
        return from_dev_cm(out), viol.value

    def hyb_delta(self, gloc, sigma, h_imp, orbit, beta, mu):
        """Construct_hyb_delta (Construct_hyb_delta.cpp:47-119): Delta(iw_n) of the
        impurity whose orbitals are `orbit` (indices into the correlated block)."""
        g, s = to_dev_cm(gloc), to_dev_cm(sigma)
        hi = to_dev_cm(np.asarray(h_imp)[None])[0].contiguous()
        nf, nc = g.shape[0], g.shape[1]
        orbit = np.ascontiguousarray(np.asarray(orbit, dtype=np.int32))
        dim = orbit.size
        out = torch.zeros((nf, max(dim, 1), max(dim, 1)), dtype=torch.complex128, device=g.device)
        self.check(self.lib.dmftgpu_hyb_delta(self.h, nc, nf, beta, mu, _ptr(g), _ptr(s), _ptr(hi), dim,
                                              orbit.ctypes.data_as(C.POINTER(C.c_int)), _ptr(out)))
        return from_dev_cm(out)[:, :dim, :dim]

    def hk_from_hr(self, n, hops, kphase):
        """H(k) = sum_R H(R) exp(-i k.a.R) (ConstructModelHamiltonian.cpp:58-77) from a hopping
        list [(R1, R2, R3, i, j, t), ...] (0-based orbitals); kphase [nk][3] = k.a per axis.
        Returns the column-major device tensor [nk][n][n] (the layout LatticeGF takes)."""
        idx = np.ascontiguousarray(np.array([h[:5] for h in hops], dtype=np.int32).reshape(-1, 5))
        t = np.ascontiguousarray(np.array([complex(h[5]) for h in hops], dtype=np.complex128))
        kp = torch.from_numpy(np.ascontiguousarray(np.asarray(kphase, dtype=np.float64))).to("cuda")
        nk = kp.shape[0]
        out = torch.empty((nk, n, n), dtype=torch.complex128, device="cuda")
        self.check(self.lib.dmftgpu_hk_from_hr(self.h, n, len(hops), idx.ctypes.data_as(C.POINTER(C.c_int)),
                                               t.ctypes.data_as(C.POINTER(C.c_double)), nk, _ptr(kp), _ptr(out)))
        out._dmft_colmajor = True
        return out

    def upfold_density(self, dm_model, corr_base, ks_energy, val_to_ks, transform, ks_ortho, overlap, beta, mu,
                       r_idx, kphase, n_k_global=None, want_dm_k=True):
        """upfolding_density (upfolding.cpp:38-139): returns (D [nk][n_orb][n_orb] or None,
        dmr [3][n_r]) with dmr = dual-dual, dual-direct, direct-direct DM(R); overlap None means S = I."""
        X = to_dev_cm(dm_model)
        nk, n = X.shape[0], X.shape[1]
        eps = torch.from_numpy(np.ascontiguousarray(np.asarray(ks_energy, dtype=np.float64))).to("cuda")
        n_orb = eps.shape[1]
        v = np.ascontiguousarray(np.asarray(val_to_ks, dtype=np.int32))
        ri = np.ascontiguousarray(np.asarray(r_idx, dtype=np.int32).reshape(-1, 5))
        kp = torch.from_numpy(np.ascontiguousarray(np.asarray(kphase, dtype=np.float64))).to("cuda")
        c, t, ko = to_dev_cm(corr_base), to_dev_cm(transform), to_dev_cm(ks_ortho)
        s = None if overlap is None else to_dev_cm(overlap)
        dm = torch.empty((nk, n_orb, n_orb), dtype=torch.complex128, device="cuda") if want_dm_k else None
        n_r = ri.shape[0]
        dmr = torch.empty((3, max(n_r, 1)), dtype=torch.complex128, device="cuda")
        self.check(self.lib.dmftgpu_upfold_density(
            self.h, n, n_orb, nk, n_k_global or nk, beta, mu, _ptr(X), _ptr(c), _ptr(eps),
            v.ctypes.data_as(C.POINTER(C.c_int)), _ptr(t), _ptr(ko), None if s is None else _ptr(s), n_r,
            ri.ctypes.data_as(C.POINTER(C.c_int)), _ptr(kp), None if dm is None else _ptr(dm),
            _ptr(dmr) if n_r else None))
        return (None if dm is None else from_dev_cm(dm)), dmr[:, :n_r].cpu().numpy()


class LatticeGF(Handle):
    """One lattice (H(k) + Sigma tail moments) bound to a library handle."""

    def __init__(self, Hk, n_corr, n_freq, beta, moments, n_cluster=1, n_k_global=None, high_freq=2, stream=None):
        super().__init__(stream)
        self.Hk = Hk if isinstance(Hk, torch.Tensor) and Hk.is_cuda and Hk.dtype == torch.complex128 \
            and getattr(Hk, "_dmft_colmajor", False) else to_dev_cm(Hk)
        nk, n = self.Hk.shape[0], self.Hk.shape[1]
        self.n, self.nc, self.nk = n, n_corr, nk
        self.n_freq, self.beta = n_freq, beta
        self.lat = capi.Lattice(n, n_corr, n_cluster, nk, n_k_global or nk, n_freq, high_freq, beta)
        self.mom = to_dev_cm(moments)
        self.check(self.lib.dmftgpu_prepare(self.h, C.byref(self.lat), _ptr(self.Hk), _ptr(self.mom)))

    def set_moments(self, moments):
        """New Sigma tail moments, same H(k): dmftgpu_update_moments."""
        self.mom = to_dev_cm(moments)
        self.check(self.lib.dmftgpu_update_moments(self.h, _ptr(self.mom)))

    def sigma_dev(self, sigma):
        if isinstance(sigma, torch.Tensor) and getattr(sigma, "_dmft_colmajor", False):
            return sigma
        s = to_dev_cm(sigma)
        s._dmft_colmajor = True
        return s

    def set_core_bands(self, ks_energy, val_to_ks):
        """KS_eigenEnergy [nk][NumOrbit] and FromValToKS [nk][n] (-1 = padded band) for the
        core-band term of GiwToNele (GreenFtn.cpp:622-631); None removes them."""
        if ks_energy is None:
            self.check(self.lib.dmftgpu_set_core_bands(self.h, 0, None, None))
            return
        eps = np.ascontiguousarray(np.asarray(ks_energy, dtype=np.float64))
        v2k = np.ascontiguousarray(np.asarray(val_to_ks, dtype=np.int32))
        self.check(self.lib.dmftgpu_set_core_bands(self.h, eps.shape[1], eps.ctypes.data_as(C.c_void_p),
                                                   v2k.ctypes.data_as(C.POINTER(C.c_int))))

    def static_eig(self):
        ev = torch.zeros((self.nk, self.n), dtype=torch.float64, device="cuda")
        U = torch.zeros((self.nk, self.n, self.n), dtype=torch.complex128, device="cuda")
        self.check(self.lib.dmftgpu_static_eig(self.h, _ptr(ev), _ptr(U)))
        torch.cuda.synchronize()
        return ev.cpu().numpy(), from_dev_cm(U)

    def greenftn_w_dev(self, sigma_dev, mu, method=PLAIN_F64, gloc=None, num=None, gloc_hf=None, want_n=False):
        nc = self.nc
        if gloc is None:
            gloc = torch.empty((self.n_freq, nc, nc), dtype=torch.complex128, device="cuda")
        if num is None:
            num = torch.empty((nc, nc), dtype=torch.complex128, device="cuda")
        N = C.c_double(0)
        self.check(self.lib.dmftgpu_greenftn_w(self.h, method, _ptr(sigma_dev), mu, _ptr(gloc), _ptr(gloc_hf),
                                               _ptr(num), C.byref(N) if want_n else None))
        if want_n:
            return gloc, num, N.value
        return gloc, num

    def greenftn_w(self, sigma, mu, method=PLAIN_F64, want_hf=False):
        """GreenFtn_w (GreenFtn.cpp:142-422) -> dict(gloc, num_matrix, n_total[, gloc_hf]) in math layout."""
        s = self.sigma_dev(sigma)
        hf = torch.empty((self.n_freq, self.nc, self.nc), dtype=torch.complex128, device="cuda") if want_hf else None
        gl, num, N = self.greenftn_w_dev(s, mu, method, gloc_hf=hf, want_n=True)
        out = dict(gloc=from_dev_cm(gl), num_matrix=from_dev_cm(num), n_total=N)
        if want_hf:
            out["gloc_hf"] = from_dev_cm(hf)
        return out

    def density_matrix(self, sigma, mu):
        s = self.sigma_dev(sigma)
        out = torch.empty((self.nk, self.n, self.n), dtype=torch.complex128, device="cuda")
        self.check(self.lib.dmftgpu_density_matrix(self.h, _ptr(s), mu, _ptr(out)))
        return from_dev_cm(out)

    def greenftn_real(self, sigma_r, mu, e0, de, eta, method=PLAIN_F64):
        """retarded_GreenFtn2 (GreenFtn.cpp:425-466) summed over k on E_j = e0 + j de:
        returns (G^R_loc [n_e][nc][nc], (1/Nk) sum_k tr G^R [n_e])."""
        s = to_dev_cm(sigma_r)
        ne = s.shape[0]
        g = torch.empty((ne, self.nc, self.nc), dtype=torch.complex128, device="cuda")
        tr = torch.empty((ne,), dtype=torch.complex128, device="cuda")
        self.check(self.lib.dmftgpu_greenftn_real(self.h, method, ne, e0, de, eta, _ptr(s), mu, _ptr(g), _ptr(tr)))
        return from_dev_cm(g), tr.cpu().numpy()

    def greenftn_real_k(self, sigma_r, mu, e0, de, eta, method=PLAIN_F64):
        """Per-k retarded G (the band spectra of qs_spectra.cpp:197-221): returns
        (G^R_cc [n_k][n_e][nc][nc], tr G^R [n_k][n_e]) with no k average."""
        s = to_dev_cm(sigma_r)
        ne = s.shape[0]
        g = torch.empty((self.nk, ne, self.nc, self.nc), dtype=torch.complex128, device="cuda")
        tr = torch.empty((self.nk, ne), dtype=torch.complex128, device="cuda")
        self.check(self.lib.dmftgpu_greenftn_real_k(self.h, method, ne, e0, de, eta, _ptr(s), mu, _ptr(g), _ptr(tr)))
        return from_dev_cm(g), tr.cpu().numpy()

    def density(self, sigma, mu, method=PLAIN_F64):
        """(N(mu), dN/dmu) -- GiwToNele (GreenFtn.cpp:513-635) plus its derivative."""
        s = self.sigma_dev(sigma)
        N, dN = C.c_double(0), C.c_double(0)
        self.check(self.lib.dmftgpu_density(self.h, method, _ptr(s), mu, C.byref(N), C.byref(dN)))
        return N.value, dN.value

    def adjust_mu(self, sigma, mu0, n_target, method=PLAIN_F64, tol=1e-10):
        """GreenFtn_w_adjust_mu (GreenFtn.cpp:66-139), Newton on the analytic dN/dmu."""
        s = self.sigma_dev(sigma)
        mu, ne = C.c_double(0), C.c_int(0)
        self.check(self.lib.dmftgpu_adjust_mu(self.h, method, _ptr(s), mu0, n_target, tol, C.byref(mu), C.byref(ne)))
        return mu.value, ne.value

    def vjp_dev(self, sigma_dev, mu, gbar_dev, nbar, method=PLAIN_F64, out=None):
        if out is None:
            out = torch.empty_like(sigma_dev)
        mb = C.c_double(0)
        self.check(self.lib.dmftgpu_gloc_vjp(self.h, method, _ptr(sigma_dev), mu, _ptr(gbar_dev), nbar, _ptr(out),
                                             C.byref(mb)))
        return out, mb.value

    def loss_grad_dev(self, sigma_dev, mu, gtgt_dev, n_target, method=PLAIN_F64, gloc=None, num=None, out=None):
        """dmftgpu_gloc_loss_grad: L = sum ||G_loc - G_tgt||^2 + (N - n_target)^2 and (dL/dSigma, dL/dmu).
        Returns (gloc, num_matrix, N, L, sigma_bar, mu_bar); tensors stay on the device."""
        gloc = torch.empty_like(gtgt_dev) if gloc is None else gloc
        num = torch.empty((self.nc, self.nc), dtype=torch.complex128, device="cuda") if num is None else num
        out = torch.empty_like(sigma_dev) if out is None else out
        N, loss, mb = C.c_double(0), C.c_double(0), C.c_double(0)
        self.check(self.lib.dmftgpu_gloc_loss_grad(self.h, method, _ptr(sigma_dev), mu, _ptr(gtgt_dev), n_target,
                                                   _ptr(gloc), _ptr(num), C.byref(N), C.byref(loss), _ptr(out),
                                                   C.byref(mb)))
        return gloc, num, N.value, loss.value, out, mb.value

    def loss_grad(self, sigma, mu, gtgt, n_target, method=PLAIN_F64):
        gl, num, N, loss, sb, mb = self.loss_grad_dev(self.sigma_dev(sigma), mu, to_dev_cm(gtgt), n_target, method)
        return dict(gloc=from_dev_cm(gl), num_matrix=from_dev_cm(num), n_total=N, loss=loss,
                    sigma_bar=from_dev_cm(sb), mu_bar=mb)

    def jvp_dev(self, sigma_dev, mu, dsig_dev, dmu, method=PLAIN_F64, out=None):
        if out is None:
            out = torch.empty_like(sigma_dev)
        dn = C.c_double(0)
        self.check(self.lib.dmftgpu_gloc_jvp(self.h, method, _ptr(sigma_dev), mu, _ptr(dsig_dev), dmu, _ptr(out),
                                             C.byref(dn)))
        return out, dn.value

    def jvp(self, sigma, mu, dsigma, dmu, method=PLAIN_F64):
        """(dG_loc, dN) for the tangent (dsigma, dmu): Jacobian-vector product, moments fixed."""
        out, dn = self.jvp_dev(self.sigma_dev(sigma), mu, to_dev_cm(dsigma), dmu, method)
        return from_dev_cm(out), dn

    def vjp(self, sigma, mu, gbar, nbar, method=PLAIN_F64):
        """(sigma_bar, mu_bar) for cotangents (gbar of G_loc, nbar of N); dL/dRe + i dL/dIm."""
        s = self.sigma_dev(sigma)
        g = to_dev_cm(gbar)
        sb, mb = self.vjp_dev(s, mu, g, nbar, method)
        return from_dev_cm(sb), mb


class Batch:
    """Many lattices in one replayable step (dmftgpu_batch_*).

    systems: list of dicts with Hk, sigma, moments (math layout), beta, mu,
    n_cluster (optional) and, for the loss + gradient step, gloc_target and
    n_target.  Outputs after run(): gloc, num_matrix, n_total and (with
    targets) sigma_bar, mu_bar per system."""

    def __init__(self, handle, systems, method=SCHUR_F64, tune=True, graph=True, prepare=False):
        self.handle = handle
        self.keep = []
        descs = []
        self.out = []
        for sd in systems:
            Hk = to_dev_cm(sd["Hk"])
            sig = to_dev_cm(sd["sigma"])
            mom = to_dev_cm(sd["moments"])
            nk, n = Hk.shape[0], Hk.shape[1]
            nf, nc = sig.shape[0], sig.shape[1]
            o = dict(gloc=torch.zeros((nf, nc, nc), dtype=torch.complex128, device="cuda"),
                     num=torch.zeros((nc, nc), dtype=torch.complex128, device="cuda"),
                     N=torch.zeros(1, dtype=torch.float64, device="cuda"))
            tgt = sd.get("gloc_target")
            if tgt is not None:
                o["tgt"] = to_dev_cm(tgt)
                o["sb"] = torch.zeros_like(sig)
                o["mb"] = torch.zeros(1, dtype=torch.float64, device="cuda")
            lat = capi.Lattice(n, nc, sd.get("n_cluster", 1), nk, nk, nf, 2, sd["beta"])
            d = capi.System(lat, Hk.data_ptr(), sig.data_ptr(), mom.data_ptr(), sd["mu"], o["gloc"].data_ptr(),
                            o["num"].data_ptr(), o["N"].data_ptr(),
                            o["tgt"].data_ptr() if tgt is not None else None, sd.get("n_target", 0.0),
                            o["sb"].data_ptr() if tgt is not None else None,
                            o["mb"].data_ptr() if tgt is not None else None)
            descs.append(d)
            self.keep += [Hk, sig, mom]
            self.out.append(o)
        arr = (capi.System * len(descs))(*descs)
        flags = (capi.BATCH_TUNE if tune else 0) | (capi.BATCH_GRAPH if graph else 0) | \
            (capi.BATCH_PREPARE if prepare else 0)
        self.b = C.c_void_p()
        handle.check(handle.lib.dmftgpu_batch_create(handle.h, len(descs), arr, method, flags, C.byref(self.b)))

    def run(self):
        self.handle.check(self.handle.lib.dmftgpu_batch_run(self.b))

    def results(self):
        torch.cuda.synchronize()
        res = []
        for o in self.out:
            r = dict(gloc=from_dev_cm(o["gloc"]), num_matrix=from_dev_cm(o["num"]), n_total=float(o["N"].item()))
            if "sb" in o:
                r["sigma_bar"] = from_dev_cm(o["sb"])
                r["mu_bar"] = float(o["mb"].item())
            res.append(r)
        return res

    def tuning(self):
        out = []
        g = 0
        while True:
            nc, ns = C.c_int(0), C.c_int(0)
            wg = (C.c_int * 2)()
            tg = (C.c_double * 2)()
            ms = (C.c_double * 2)()
            rc = self.handle.lib.dmftgpu_batch_tuning(self.b, g, C.byref(nc), C.byref(ns), wg, tg, ms)
            if rc != 0:
                break
            out.append(dict(n_corr=nc.value, systems=ns.value, wg=list(wg), target=list(tg), ms=list(ms)))
            g += 1
        return out

    def close(self):
        if self.b:
            self.handle.lib.dmftgpu_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
