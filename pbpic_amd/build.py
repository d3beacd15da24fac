"""Builds libpbpic.so (hipcc, gfx950) in the package directory.

Run as ``python -m pbpic_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpbpic.so")

SOURCES = ["capi.hip", "k_push.hip", "k_field.hip", "k_sort.hip", "k_adjoint.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=True, out=None):
    """One hipcc per translation unit, in parallel, then one link.  The
    translation units share no device symbols, so each carries its own code
    object."""
    out = out or LIB
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    procs, objs = [], []
    for s in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc(), *compile_flags, "-I", os.path.join(ROOT, "include"), "-c",
               os.path.join(CSRC, s), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd)))
    failed = [s for s, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, f"hipcc {' '.join(failed)}")
    cmd = [hipcc(), "--offload-arch=gfx950", "-fPIC", "-shared", *objs, "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


def build_fortran(verbose=True):
    """Fortran driver mirroring `program pic2d`, linked against libpbpic.so."""
    fdir = os.path.join(HERE, "fortran")
    fc = "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        return None
    out = os.path.join(fdir, "pic2d_gpu")
    cmd = [fc, "-O2", "-J", fdir, os.path.join(fdir, "pbpic_mod.f90"),
           os.path.join(fdir, "pic2d_driver.f90"), "-o", out,
           "-L" + HERE, "-lpbpic", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    # the test program that moves the reference's own host state (type
    # particle / fields / density arrays) through the module, in the
    # reference's default 4-byte real and in a -fdefault-real-8 build
    tdir = os.path.join(ROOT, "tests", "fortran")
    for kind, flags in (("r4", []), ("r8", ["-fdefault-real-8"])):
        mdir = os.path.join(HERE, "_build", "fortran_" + kind)
        os.makedirs(mdir, exist_ok=True)
        cmd = [fc, "-O2", *flags, "-J", mdir, os.path.join(fdir, "pbpic_mod.f90"),
               os.path.join(tdir, "ref_state_step.f90"), "-o",
               os.path.join(tdir, "ref_state_step_" + kind), "-L" + HERE, "-lpbpic",
               "-Wl,-rpath," + os.path.relpath(HERE, tdir).join(["$ORIGIN/", ""])]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=mdir)
    return out


if __name__ == "__main__":
    build_library()
    if len(sys.argv) > 1 and sys.argv[1] == "--fortran":
        build_fortran()
