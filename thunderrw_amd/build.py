"""Builds thunderrw_amd/libthunderrw.so (HIP kernels for gfx950 + the C ABI).

    python -m thunderrw_amd.build        (or: python thunderrw_amd/build.py)

Objects go to thunderrw_amd/_build/, the shared library next to this file.
Every HIP translation unit is compiled with contraction off (the FP64 table
and sampling arithmetic must round exactly as written) and with packed FP32
instructions disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libthunderrw.so")
ARCH = os.environ.get("TRW_OFFLOAD_ARCH", "gfx950")

SOURCES = ["walk.hip", "tables.hip", "probe.hip", "serve.hip", "api.cpp", "server.cpp", "autotune.cpp", "sfmt.cpp",
           "hostgraph.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), "-Wall"]
HIP_ONLY = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    flags = COMMON + HIP_ONLY
    cmd = [hipcc()] + flags + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


DROPIN = os.path.join(HERE, "dropin")
DROPIN_APPS = ["deepwalk", "ppr", "node2vec", "metapath", "genericwalk"]
# Apps of this repository's tests whose functors have no built-in Move (functor engine).
TEST_APPS_DIR = os.path.join(ROOT, "tests", "dropin_apps")


def compile_app(src, exe):
    """One app translation unit against dropin/creeper_main.h and this library, with hipcc
    (the functor engine instantiates kernels for the app's functor)."""
    cmd = [hipcc(), "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=" + ARCH, "-Xclang",
           "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-I", DROPIN,
           "-I", os.path.join(ROOT, "include"), src, "-o", exe, "-L", HERE, "-lthunderrw",
           "-Wl,-rpath,$ORIGIN/../.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in build failed: %s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    return exe


def build_dropin(reference_dir, verbose=False):
    """Compiles the reference's own app sources, unchanged, against dropin/creeper_main.h and this
    library, into dropin/bin/<app> (a checkout of the reference is needed; the binaries are not
    committed)."""
    apps_dir = os.path.join(reference_dir, "random_walk", "apps")
    if not os.path.isdir(apps_dir):
        return []
    out_dir = os.path.join(DROPIN, "bin")
    os.makedirs(out_dir, exist_ok=True)
    jobs = [(os.path.join(apps_dir, app + ".cpp"), os.path.join(out_dir, app)) for app in DROPIN_APPS]
    with cf.ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        exes = list(ex.map(lambda j: compile_app(*j), jobs))
    if verbose:
        print("built drop-in apps:", " ".join(DROPIN_APPS))
    return exes


def build_test_apps(verbose=False):
    """tests/dropin_apps/<name>.cpp -> dropin/bin/<name> (no reference checkout needed)."""
    if not os.path.isdir(TEST_APPS_DIR):
        return []
    out_dir = os.path.join(DROPIN, "bin")
    os.makedirs(out_dir, exist_ok=True)
    srcs = sorted(f for f in os.listdir(TEST_APPS_DIR) if f.endswith(".cpp"))
    jobs = [(os.path.join(TEST_APPS_DIR, f), os.path.join(out_dir, f[:-4])) for f in srcs]
    with cf.ThreadPoolExecutor(max_workers=max(1, len(jobs))) as ex:
        exes = list(ex.map(lambda j: compile_app(*j), jobs))
    if verbose and exes:
        print("built test apps:", " ".join(os.path.basename(e) for e in exes))
    return exes


if __name__ == "__main__":
    build(verbose=True)
    build_test_apps(verbose=True)
    if len(sys.argv) > 1:
        build_dropin(sys.argv[1], verbose=True)
    sys.exit(0)
