"""dmftgpu_greenftn_w_noalloc through the C ABI against dmftgpu_greenftn_w.

The allocation-free forward takes its scratch from a fixed device array and
may chunk k more coarsely (the partial sums must fit), so its sums run in a
different order: the bar is the FP64 one, 1e-12 normwise relative (FP32 +
refinement 1e-9), on G_loc, GwHF, NumMatrix and N.  Two calls in a row, the
second on a smaller lattice (the array keeps the first call's contents), and
a lattice too large for the array, which must be refused.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from dmftpack_amd import capi, lattice as L
from oracle import oracle as O


def rel(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / max(np.max(np.abs(ref)), 1e-300))


def both(lat, sig, mu, method):
    lib = capi.lib()
    from dmftpack_amd import gpu as G
    s = lat.sigma_dev(sig)
    out = {}
    for name, fn in (("alloc", lib.dmftgpu_greenftn_w), ("noalloc", lib.dmftgpu_greenftn_w_noalloc)):
        gl = torch.full((lat.n_freq, lat.nc, lat.nc), float("nan"), dtype=torch.complex128, device="cuda")
        hf = torch.full_like(gl, float("nan"))
        nm = torch.full((lat.nc, lat.nc), float("nan"), dtype=torch.complex128, device="cuda")
        N = C.c_double(0)
        lat.check(fn(lat.h, method, G._ptr(s), mu, G._ptr(gl), G._ptr(hf), G._ptr(nm), C.byref(N)))
        torch.cuda.synchronize()
        out[name] = dict(gloc=G.from_dev_cm(gl), hf=G.from_dev_cm(hf), num=G.from_dev_cm(nm), N=N.value)
    return out["alloc"], out["noalloc"]


@pytest.mark.gpu
def test_noalloc_matches_greenftn_w():
    from dmftpack_amd import gpu as G
    # first call: 16^3 k points, 1024 frequencies (the scratch holds fewer k chunks
    # than the default plan); second call on the same handle: a smaller lattice
    w = L.workload("dp14", seed=3, nk_side=16, n_freq=1024)
    mom = w["exact_moments"]
    lat = G.LatticeGF(w["Hk"], 5, 1024, w["beta"], mom)
    for method, tol in ((G.SCHUR_F64, 1e-12), (G.PLAIN_F64, 1e-12), (G.SCHUR_F32R, 1e-9)):
        a, b = both(lat, w["sigma"], 0.2, method)
        print(f"no-alloc 16^3 x 1024 method {method}: G_loc {rel(b['gloc'], a['gloc']):.1e} "
              f"NumMatrix {rel(b['num'], a['num']):.1e} N {abs(b['N'] - a['N']):.1e}", flush=True)
        assert rel(b["gloc"], a["gloc"]) <= tol and rel(b["hf"], a["hf"]) <= 1e-12
        assert rel(b["num"], a["num"]) <= tol and abs(b["N"] - a["N"]) <= tol * abs(a["N"])
    Hk = L.random_tb_hk(7, 3, 5)
    sig, mom = L.pole_sigma(3, 48, 30.0, 5)
    # same handle, new (smaller) lattice bound by dmftgpu_prepare
    lat.Hk = G.to_dev_cm(Hk)
    lat.n, lat.nc, lat.nk, lat.n_freq, lat.beta = 7, 3, Hk.shape[0], 48, 30.0
    lat.lat = capi.Lattice(7, 3, 1, Hk.shape[0], Hk.shape[0], 48, 2, 30.0)
    lat.mom = G.to_dev_cm(mom)
    lat.check(lat.lib.dmftgpu_prepare(lat.h, C.byref(lat.lat), G._ptr(lat.Hk), G._ptr(lat.mom)))
    ref = O.greenftn_w(Hk, sig, mom, 0.1, 30.0)
    for method, tol in ((G.SCHUR_F64, 1e-12), (G.PLAIN_F64, 1e-12), (G.SCHUR_F32R, 1e-9)):
        a, b = both(lat, sig, 0.1, method)
        print(f"no-alloc small method {method}: G_loc {rel(b['gloc'], a['gloc']):.1e}", flush=True)
        assert rel(b["gloc"], a["gloc"]) <= tol and rel(b["num"], a["num"]) <= tol
        assert rel(b["gloc"], ref["gloc"]) <= tol and rel(b["hf"], ref["gloc_hf"]) <= 1e-12
    lat.close()


@pytest.mark.gpu
def test_noalloc_rejects_lattice_beyond_scratch():
    from dmftpack_amd import gpu as G
    # n_corr = 16 at 4096 frequencies (8192 summed): one k chunk alone needs more than 128 MiB
    Hk = L.random_tb_hk(17, 1, 9)
    sig, mom = L.pole_sigma(16, 4096, 100.0, 9)
    lat = G.LatticeGF(Hk, 16, 4096, 100.0, mom)
    s = lat.sigma_dev(sig)
    gl = torch.zeros((4096, 16, 16), dtype=torch.complex128, device="cuda")
    nm = torch.zeros((16, 16), dtype=torch.complex128, device="cuda")
    rc = capi.lib().dmftgpu_greenftn_w_noalloc(lat.h, G.SCHUR_F64, G._ptr(s), 0.0, G._ptr(gl), None, G._ptr(nm), None)
    assert rc == capi.ERR_UNSUPPORTED
    assert b"scratch" in capi.lib().dmftgpu_last_error(lat.h)
    # the allocating entry point still handles it
    lat.check(capi.lib().dmftgpu_greenftn_w(lat.h, G.SCHUR_F64, G._ptr(s), 0.0, G._ptr(gl), None, G._ptr(nm), None))
    lat.close()
