"""In-tree build of the HIP library (gfx950) and the checker.

    python -m mhd2dcfdfftw_open_amd.build          # both libraries
Outputs (git-ignored, shipped with the tree):
    mhd2dcfdfftw_open_amd/libmhd2d.so       kernels + context API + drop-in
                                            entry points + host storage of the
                                            reference globals (standalone)
    mhd2dcfdfftw_open_amd/libmhd2d_core.so  the same without the globals, for
                                            linking against the reference's own
                                            initialize.o (globals resolved from
                                            the executable)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.environ.get("MHD2D_BUILD_DIR", os.path.join(PKG, "_build"))
LIB = os.environ.get("MHD2D_LIB_OUT", os.path.join(PKG, "libmhd2d.so"))
LIB_CORE = os.path.join(os.path.dirname(LIB), os.path.basename(LIB).replace(
    "libmhd2d", "libmhd2d_core"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# extra compile flags (e.g. tuning -D switches of the kernels) for variant builds
EXTRA = os.environ.get("MHD2D_EXTRA_FLAGS", "").split()
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", *EXTRA,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]
CORE_SOURCES = ["mhd2d_kernels.hip", "mhd2d_xgemm.hip", "mhd2d_tables.cpp", "mhd2d_ctx.cpp",
                "mhd2d_dropin.cpp", "mhd2d_ic.cpp", "mhd2d_io.cpp"]
# the storage of the reference globals and what needs it (standalone library)
GLOBAL_SOURCES = ["mhd2d_globals.cpp", "mhd2d_driver.cpp"]
RUN_EXE = os.path.join(os.path.dirname(LIB), "mhd2d_run")


def _compile(src):
    out = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", *COMMON, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msgs}")
    return out, msgs


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = CORE_SOURCES + GLOBAL_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = dict(zip(srcs, ex.map(_compile, srcs)))
    for s, (_, msgs) in results.items():
        if msgs and verbose:
            print(f"[{s}]\n{msgs}")
    core_objs = [results[s][0] for s in CORE_SOURCES]
    glob_objs = [results[s][0] for s in GLOBAL_SOURCES]
    for out, objs in ((LIB, core_objs + glob_objs), (LIB_CORE, core_objs)):
        tmp = out + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed for {out}:\n{r.stdout}{r.stderr}")
        os.replace(tmp, out)
    # the reference-style executable, linked against the standalone library
    tmp = RUN_EXE + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "mhd2d_run_main.cpp"), "-o", tmp, LIB,
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed for {RUN_EXE}:\n{r.stdout}{r.stderr}")
    os.replace(tmp, RUN_EXE)
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    for p in (LIB, LIB_CORE, RUN_EXE):
        if os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    if "clean" in sys.argv[1:]:
        clean()
    print(build(verbose=True))
