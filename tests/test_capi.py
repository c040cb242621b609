"""The C-ABI library loads on a machine without a GPU and exports every entry point that
include/glia_rd.h, include/glia_io.h, include/glia_opt.h, include/glia_adv.h, include/glia_el.h,
include/glia_ms.h and include/glia_me.h declare (no compute calls here)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header="glia_rd.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(glia_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("glia_rd_create", "glia_rd_destroy", "glia_rd_solve_state", "glia_rd_solve_adjoint",
                 "glia_rd_objective", "glia_rd_gradient", "glia_rd_objective_and_gradient",
                 "glia_rd_set_tissue", "glia_rd_set_basis", "glia_rd_set_data", "glia_rd_stats",
                 "glia_rd_last_error", "glia_rd_diffusion_solve"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from glia_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(glia_[a-z0-9_]+)\b", out))
    assert set(declared()) <= exported
    assert set(declared("glia_io.h")) <= exported
    assert set(declared("glia_opt.h")) <= exported
    assert set(declared("glia_adv.h")) <= exported
    assert set(declared("glia_el.h")) <= exported
    assert set(declared("glia_ms.h")) <= exported
    assert set(declared("glia_me.h")) <= exported


def test_python_binding_covers_every_symbol():
    from glia_amd import _lib
    assert set(declared()) == set(_lib.SIGNATURES)
    assert set(declared("glia_io.h")) == set(_lib.IO_SIGNATURES)
    assert set(declared("glia_opt.h")) == set(_lib.OPT_SIGNATURES)
    assert set(declared("glia_adv.h")) == set(_lib.ADV_SIGNATURES)
    assert set(declared("glia_el.h")) == set(_lib.EL_SIGNATURES)
    assert set(declared("glia_ms.h")) == set(_lib.MS_SIGNATURES)
    assert set(declared("glia_me.h")) == set(_lib.ME_SIGNATURES)


def test_config_defaults_without_gpu():
    from glia_amd import _lib
    lib = _lib.load()
    c = _lib.Config()
    lib.glia_rd_config_init(ctypes.byref(c))
    assert list(c.n) == [64, 64, 64] and c.dt == 0.5 and c.nt == 1 and c.order == 2
    assert c.pcg_rtol == 1e-6 and c.pcg_maxit == 5000 and c.k_scale == 1.0


def test_create_rejects_bad_arguments_without_touching_the_gpu():
    from glia_amd import _lib
    lib = _lib.load()
    c = _lib.Config()
    lib.glia_rd_config_init(ctypes.byref(c))
    c.order = 1
    h = ctypes.c_void_p()
    assert lib.glia_rd_create(ctypes.byref(c), ctypes.byref(h)) == _lib.ERR_UNSUPPORTED
    c.order = 2
    c.n[2] = 48  # the packed half-spectrum row (24) is not a multiple of the 16-column tile
    assert lib.glia_rd_create(ctypes.byref(c), ctypes.byref(h)) == _lib.ERR_UNSUPPORTED
    assert lib.glia_rd_create(None, ctypes.byref(h)) == _lib.ERR_ARG
    assert lib.glia_rd_destroy(None) == _lib.ERR_ARG


def test_cpp_interface_headers_compile():
    """include/glia/ (the reference's class names over the C ABI) is self-contained C++17."""
    src = '#include "glia/glia.h"\nint main() { return 0; }\n'
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

