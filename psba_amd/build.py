"""Build the HIP library (libpsba.so, gfx950) and the CPU oracle.

    python -m psba_amd.build            # everything
    python -m psba_amd.build --lib      # product library only

Objects go to psba_amd/_build/; the shared libraries are written in-tree
(psba_amd/libpsba.so, oracle/liboracle.so) so they travel with the tree.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "psba_amd", "csrc")
BUILD = os.path.join(ROOT, "psba_amd", "_build")
LIB = os.path.join(ROOT, "psba_amd", "libpsba.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

ARCH = os.environ.get("PSBA_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

SOURCES = ["ctx.cpp", "drivers.cpp", "dropin.cpp", "scene.cpp", "sba_io.cpp", "batch.cpp", "server.cpp",
           "k_obs.hip", "k_schur.hip", "k_dense.hip", "k_tr.hip", "k_batch.hip", "k_index.hip", "k_lm.hip", "k_noalloc.hip"]

HIP_FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Wall", "-Wno-unused-function",
    # packed FP32 instructions stay off (they are produced only by the compiler)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


STAMP = os.path.join(ROOT, "psba_amd", ".build_stamp")


def source_hash():
    """SHA-256 over every source the built files come from (the HIP/C++
    library, its headers, the oracle, the drop-in program, this script)."""
    import hashlib
    h = hashlib.sha256()
    files = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC))]
    inc = os.path.join(ROOT, "include")
    files += [os.path.join(inc, f) for f in sorted(os.listdir(inc))]
    files += [os.path.join(ORACLE_DIR, "oracle.c"), os.path.join(ORACLE_DIR, "oracle.h"),
              os.path.join(ROOT, "tests", "dropin", "main_dropin.cpp"), os.path.abspath(__file__)]
    for f in files:
        if os.path.isfile(f):
            h.update(os.path.relpath(f, ROOT).encode())
            with open(f, "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def built_hash():
    """The source hash recorded by the last complete build, or None."""
    try:
        with open(STAMP) as f:
            return f.read().strip()
    except OSError:
        return None


def is_current():
    outs = (LIB, ORACLE_LIB, DROPIN_BIN)
    return all(os.path.exists(p) for p in outs) and built_hash() == source_hash()


def build_all(jobs=8, verbose=False):
    """Everything, then the stamp of the sources it was built from."""
    h = source_hash()
    build_oracle()
    build_lib(jobs, verbose)
    build_dropin()
    with open(STAMP, "w") as f:
        f.write(h + "\n")
    return h


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_lib(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


DROPIN_SRC = os.path.join(ROOT, "tests", "dropin", "main_dropin.cpp")
DROPIN_BIN = os.path.join(ROOT, "tests", "dropin", "main_dropin")


def build_dropin():
    """The reference-style host program of tests/dropin, linked to libpsba.so."""
    tmp = DROPIN_BIN + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), DROPIN_SRC,
          "-L" + os.path.dirname(LIB), "-lpsba", "-Wl,-rpath,$ORIGIN/../../psba_amd", "-o", tmp])
    os.replace(tmp, DROPIN_BIN)
    return DROPIN_BIN


def build_oracle():
    src = os.path.join(ORACLE_DIR, "oracle.c")
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-std=c11", "-O2", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-o", tmp, src, "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    both = not (a.lib or a.oracle)
    if both:
        print("built sources", build_all(a.j, a.v))
        return
    if a.oracle:
        print("built", build_oracle())
    if a.lib:
        print("built", build_lib(a.j, a.v))
        print("built", build_dropin())


if __name__ == "__main__":
    main()
