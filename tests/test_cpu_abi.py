"""The C ABI library loads without a GPU and exports every declared entry point."""
import ctypes as C
import os
import subprocess

import pytest

from dmftpack_amd import capi


def test_library_exports_every_header_symbol():
    lib = capi.lib()
    declared = capi.header_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(capi._SIGS), "ctypes signatures must cover the header exactly"


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True, check=True)
    syms = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    for name in capi.header_symbols():
        assert name in syms, name


def test_version_and_null_handle_errors():
    lib = capi.lib()
    assert b"gfx950" in lib.dmftgpu_version()
    assert lib.dmftgpu_last_error(None) == b"null handle"
    assert lib.dmftgpu_destroy(None) == capi.ERR_ARG
    assert lib.dmftgpu_set_option(None, 1, 1) == capi.ERR_ARG
    assert lib.dmftgpu_prepare(None, None, None, None) == capi.ERR_ARG


def test_header_compiles_as_c():
    hdr = os.path.join(os.path.dirname(capi.HEADER), "dmftgpu.h")
    src = '#include "dmftgpu.h"\nint main(void){ dmftgpu_lattice l; (void)l; return DMFTGPU_OK; }\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.dirname(hdr), "-x", "c",
                        "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_lattice_struct_layout():
    assert C.sizeof(capi.Lattice) == 7 * 4 + 4 + 8  # seven int32, padding, double
    assert capi.Lattice.beta.offset == 32
