"""Build recipe for the gfx950 engine (libquantr_hip.so) and the C++ API module.

Everything is compiled in-tree into quantr_amd/lib/ so the built libraries travel with the
repository snapshot. hipcc cross-compiles for gfx950 without a GPU present.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled: hipcc must not form v_pk_*_f32 on its own.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-I" + INCLUDE, "-I" + CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
]
SOURCES = ["kernels.hip", "tile_interp.hip", "lower.cpp", "planner.cpp", "runtime.cpp", "batch.cpp", "capi.cpp", "jit.cpp", "sampler.cpp", "noalloc.cpp", "shots.cpp"]


def embed_tile_ops():
    """tile_ops.hpp as a C++ string, compiled into the library for runtime specialisation."""
    src = open(os.path.join(CSRC, "tile_ops.hpp")).read()
    assert ")QRSRC\"" not in src
    out = os.path.join(BUILD, "tile_ops_src.cpp")
    text = ("namespace qr {\nextern const char* const kTileOpsSource;\n"
            "const char* const kTileOpsSource = R\"QRSRC(" + src + ")QRSRC\";\n}\n")
    if not os.path.exists(out) or open(out).read() != text:
        with open(out, "w") as f:
            f.write(text)
    return out


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def _deps_mtime(path, seen=None):
    """Newest modification time of a source and the project headers it includes (transitively)."""
    seen = set() if seen is None else seen
    if path in seen or not os.path.exists(path):
        return 0.0
    seen.add(path)
    t = os.path.getmtime(path)
    for line in open(path, errors="replace"):
        line = line.strip()
        if line.startswith("#include \""):
            name = line.split("\"")[1]
            for d in (os.path.dirname(path), CSRC, INCLUDE):
                cand = os.path.join(d, name)
                if os.path.exists(cand):
                    t = max(t, _deps_mtime(cand, seen))
                    break
    return t


def build_engine(verbose=False, force=True):
    """Compiles every source (force=True), or only those whose object is older than the source
    or any header (force=False, for quick rebuilds while developing), and links the library."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    objs = []
    cmds = []
    embedded = embed_tile_ops()
    for src in SOURCES + [embedded]:
        path = src if os.path.isabs(src) else os.path.join(CSRC, src)
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        objs.append(obj)
        if (not force and os.path.exists(obj)
                and os.path.getmtime(obj) > _deps_mtime(path)):
            continue
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", path, "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out:
                print(out)
    so = os.path.join(LIB, "libquantr_hip.so")
    tmp = so + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lhiprtc", "-Wl,-soname,libquantr_hip.so"])
    os.replace(tmp, so)
    return so


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-s", "_build/liboracle.so"])
    return os.path.join(odir, "_build", "liboracle.so")


def build_all(verbose=False, force=True):
    so = build_engine(verbose, force)
    orc = build_oracle()
    return so, orc


if __name__ == "__main__":
    print(build_all(verbose=True, force="--changed" not in sys.argv))
