"""Build libmpdok_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Every ``csrc/*.hip`` file is compiled to an object under ``_build/`` and the
objects are linked into ``mpdok_amd/libmpdok_hip.so``. Run as
``python -m mpdok_amd.build`` or call :func:`build`.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(HERE, "libmpdok_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-mcode-object-version=5",
    # Contract a*b+c only inside one expression (C's FP_CONTRACT ON); explicit
    # __builtin_fma is used where fusion is wanted.
    "-ffp-contract=on",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-command-line-argument",
    "-I", CSRC,
]


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        ln for ln in (p.stdout + p.stderr).splitlines()
        if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln
    )
    return obj, p.returncode, out, " ".join(cmd)


def build(verbose=True, jobs=8):
    """Compile all HIP sources and link the shared library. Raises on failure."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    failed = False
    for obj, rc, out, cmd in results:
        if rc != 0:
            failed = True
            sys.stderr.write(f"FAILED: {cmd}\n{out}\n")
        elif verbose and out.strip():
            sys.stderr.write(out + "\n")
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *[r[0] for r in results], "-lpthread"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{p.stdout}{p.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


if __name__ == "__main__":
    build()
