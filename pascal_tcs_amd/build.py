"""Build libpascal_tcs.so (HIP, gfx950) in-tree with hipcc.

Every translation unit under csrc/ is compiled for gfx950 with FP contraction
off (the plain FP64 form reproduces the reference's operation order, so no
FMA fusion may change a rounding) and packed-FP32 instructions disabled, then
linked into one shared library next to this file.  The stage units
(thermal, momentum, pressure) are compiled twice more: as the `fast`
arithmetic form (contraction and reciprocals allowed; precision 1) and as
the `mixed` form (fast arithmetic, FP32 stage rows, ADI scratch and increments;
precision 2).  The Fortran drop-in
shells under fortran/ are compiled too when flang is present.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpascal_tcs.so")
SOURCES = ["runtime.hip", "tdma.hip", "tdma_part.hip", "thermal.hip", "momentum.hip", "pressure.hip", "adjoint.hip"]
# stage translation units built again as the `fast` and `mixed` forms
FORM_SOURCES = ["thermal.hip", "momentum.hip", "pressure.hip"]
ARCH = os.environ.get("PASCAL_TCS_ARCH", "gfx950")


def hipcc():
    for cand in [os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")]:
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def flags(form="exact"):
    # exact form: no contraction, IEEE division; fast form: FMA contraction,
    # reciprocals and reassociation (inf/NaN semantics kept); mixed form:
    # the fast arithmetic with FP32 rows, ADI scratch and increments
    fastm = ["-ffp-contract=fast", "-ffast-math", "-fno-finite-math-only"]
    arith = {"exact": ["-ffp-contract=off", "-fno-fast-math"],
             "fast": fastm + ["-DTCS_FORM=fast", "-DTCS_RECIP"],
             "mixed": fastm + ["-DTCS_FORM=mixed", "-DTCS_SCRATCH=float", "-DTCS_RECIP"]}[form]
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        *arith,
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # the split line solves that compute their rows unroll a loop over
        # the part's 16 rows around a whole stencil: let it unroll fully so
        # the rows stay in registers
        "-mllvm", "-pragma-unroll-threshold=60000",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    ]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False, jobs=8):
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    units = ([(s, "exact") for s in SOURCES] + [(s, "fast") for s in FORM_SOURCES]
             + [(s, "mixed") for s in FORM_SOURCES])
    for src, form in units:
        obj = os.path.join(objdir, src.replace(".hip", ".o" if form == "exact" else f"_{form}.o"))
        objs.append(obj)
        cmd = [cc] + flags(form) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _wait(procs)
            procs = []
    _wait(procs)
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_fortran(verbose=verbose)
    return LIB


def _wait(procs):
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError("compile failed: " + " ".join(cmd))


def build_fortran(verbose=False):
    """Compile the Fortran drop-in shells + driver when flang is available."""
    fdir = os.path.join(HERE, "fortran")
    mk = os.path.join(fdir, "Makefile")
    flang = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
    if not (os.path.exists(mk) and os.path.exists(flang)):
        return None
    _run(["make", "-s", "-C", fdir, f"FLANG={flang}", f"LIBDIR={HERE}"])
    return fdir


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
