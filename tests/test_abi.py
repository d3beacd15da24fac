"""CPU checks of the drop-in boundary (no GPU calls): the built library loads
and exports every entry point include/pbpic.h declares, the Fortran interface
binds the same names, and the product package never reaches the oracle."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pbpic.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbpic_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import pbpic_amd
    if not os.path.exists(pbpic_amd.LIB_PATH):
        from pbpic_amd import build
        build.build_library(verbose=False)
    return pbpic_amd.load_library()


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("pbpic_push", "pbpic_deposit", "pbpic_fieldsolve", "pbpic_countingsort",
                 "pbpic_fieldsenergy", "pbpic_step", "pbpic_diag", "pbpic_ctx_create"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_exports_are_c_symbols(lib):
    import pbpic_amd
    out = subprocess.run(["nm", "-D", "--defined-only", pbpic_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (pbpic_\w+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_fortran_interface_binds_header_names():
    src = open(os.path.join(ROOT, "pbpic_amd", "fortran", "pbpic_mod.f90")).read()
    bound = set(re.findall(r'name="(pbpic_\w+)"', src))
    assert bound and bound <= set(declared_functions())
    # the reference's subroutine names: plain subroutines, or (currentdeposit,
    # which takes real arguments of the caller's default kind) a generic
    # interface over one module procedure per real kind
    for sub in ("pusher", "fieldsolver", "countingsort", "fieldsenergy"):
        assert re.search(rf"^\s*subroutine {sub}\b", src, flags=re.M), sub
    m = re.search(r"^\s*interface currentdeposit\s*\n\s*module procedure (.+)$", src, flags=re.M)
    assert m, "currentdeposit"
    for proc in (x.strip() for x in m.group(1).split(",")):
        assert re.search(rf"^\s*subroutine {proc}\b", src, flags=re.M), proc


def test_params_struct_layout(lib):
    import pbpic_amd
    p = pbpic_amd.Params()
    lib.pbpic_params_default(ctypes.byref(p))
    # the reference's own constants (pic2d.f90:12-42)
    assert (p.nx, p.ny, p.dx, p.cva, p.mime) == (256, 256, 0.04, 15.0, 25.0)
    assert abs(p.in0 - 1 / 291) < 1e-18 and abs(p.dt - 0.045 / 25) < 1e-18
    assert p.sort_every == 50 and p.compat == 7
    assert ctypes.sizeof(pbpic_amd.Params) == 4 * 4 + 7 * 8 + 2 * 4 + 8 * 4


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "pbpic_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".f90")):
                text = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"\b(import|from)\s+oracle\b", text), f
                assert "pic_oracle" not in text and "libpbpic_oracle" not in text, f


def test_missing_library_fails_loudly(tmp_path):
    import pbpic_amd
    with pytest.raises(pbpic_amd.PbpicError):
        pbpic_amd.load_library(str(tmp_path / "absent.so"))


def test_one_hip_runtime_whatever_the_import_order():
    """Loading the library before torch must not map a second HIP runtime:
    torch bundles libamdhip64.so.7 under the same soname as ROCm's."""
    pytest.importorskip("torch")
    import pbpic_amd
    if not os.path.exists(pbpic_amd.LIB_PATH):
        pytest.skip("library not built")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import pbpic_amd; pbpic_amd.load_library()\n"
            "import torch\n"
            "m = open('/proc/self/maps').read().splitlines()\n"
            "print(len({l.split()[-1] for l in m if 'libamdhip64' in l}))\n") % ROOT
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True, check=True)
    assert out.stdout.strip().splitlines()[-1] == "1", out.stdout
