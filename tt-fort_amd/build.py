"""Build the gfx950 HIP library (and its int64-Fortran-integer twin) in-tree.

    python tt-fort_amd/build.py            # builds tt-fort_amd/lib/*.so

Every HIP source under csrc/ is compiled by hipcc for gfx950 and linked into
lib/libttfort_amd.so (Fortran INTEGER = int32) and lib/libttfort_amd_i8.so
(INTEGER = int64; only the reference-compatible entry points differ).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay disabled in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", INCLUDE, "-I", CSRC,
]

# sources compiled once; refabi.hip is compiled per Fortran integer kind
SHARED_SOURCES = ["runtime.hip", "gemm.hip", "contract.hip", "linalg.hip", "lobpcg.hip", "eigb.hip",
                  "amen_local.hip", "amen.hip", "ksl.hip", "mlp.hip",
                  "eigb_batch.hip", "sdv_io.hip", "noalloc.hip", "mvk4.hip", "amen_typed.hip",
                  "bfun3_fused.hip"]
KIND_SOURCES = ["refabi.hip"]


def _sources_present(names):
    return [n for n in names if os.path.exists(os.path.join(CSRC, n))]


def _compile(src, obj, extra):
    cmd = [HIPCC] + CFLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    return src, time.time() - t


def build(verbose: bool = True, jobs: int = 8) -> list:
    """Compile and link both library variants; returns the .so paths."""
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    shared = _sources_present(SHARED_SOURCES)
    kind = _sources_present(KIND_SOURCES)
    tasks = []
    for s in shared:
        tasks.append((s, os.path.join(OBJ, s + ".o"), []))
    for s in kind:
        tasks.append((s, os.path.join(OBJ, s + ".i4.o"), []))
        tasks.append((s, os.path.join(OBJ, s + ".i8.o"), ["-DTTF_FINT64"]))
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(tasks)))) as ex:
        futs = [ex.submit(_compile, *t) for t in tasks]
        for f in cf.as_completed(futs):
            src, dt = f.result()
            if verbose:
                print("  compiled %-20s %.1fs" % (src, dt), flush=True)
    outs = []
    for tag, suffix in (("i4", ""), ("i8", "_i8")):
        objs = [os.path.join(OBJ, s + ".o") for s in shared]
        objs += [os.path.join(OBJ, s + ".%s.o" % tag) for s in kind]
        so = os.path.join(LIB, "libttfort_amd%s.so" % suffix)
        tmp = so + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
        os.replace(tmp, so)
        outs.append(so)
    if verbose:
        print("  linked %s in %.1fs" % (", ".join(os.path.basename(o) for o in outs), time.time() - t0),
              flush=True)
    outs += build_fortran(verbose=verbose)
    return outs


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran")


def fortran_commands(kind, outdir, libdir=LIB, rpath=None):
    """Commands that compile the Fortran module and its boundary check
    program for one integer kind ('i4' or 'i8': the reference's presets,
    default INTEGER of 4 or 8 bytes) and link them against the library in
    libdir (found at run time through rpath, default libdir)."""
    flags = ["-cpp", "-O2"] + (["-DTTF_FINT64", "-fdefault-integer-8"] if kind == "i8" else [])
    lib = "ttfort_amd" if kind == "i4" else "ttfort_amd_i8"
    mod_o = os.path.join(outdir, "ttfort_amd.o")
    exe = os.path.join(outdir, "check_boundary_%s" % kind)
    return exe, [
        [FLANG] + flags + ["-c", os.path.join(FORTRAN, "ttfort_amd.F90"), "-module-dir", outdir, "-o", mod_o],
        [FLANG] + flags + ["-I", outdir, os.path.join(FORTRAN, "check_boundary.F90"), mod_o, "-L", libdir,
                           "-l" + lib, "-Wl,-rpath," + (rpath or libdir), "-o", exe],
    ]


def build_fortran(verbose=True):
    """The Fortran module and boundary check programs, per integer kind, in
    lib/fortran_<kind>/ (skipped with a message when flang is absent); the
    programs find the library relative to themselves, so the tree can move."""
    if not os.path.exists(FLANG):
        if verbose:
            print("  flang not found at %s: Fortran module not built" % FLANG, flush=True)
        return []
    outs = []
    for kind in ("i4", "i8"):
        outdir = os.path.join(LIB, "fortran_%s" % kind)
        os.makedirs(outdir, exist_ok=True)
        exe, cmds = fortran_commands(kind, outdir, rpath="$ORIGIN/..")
        for cmd in cmds:
            r = subprocess.run(cmd, capture_output=True, text=True, cwd=outdir)
            if r.returncode != 0:
                raise RuntimeError("flang failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
        outs.append(exe)
    if verbose:
        print("  built the Fortran module and boundary checks (i4, i8)", flush=True)
    return outs


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
