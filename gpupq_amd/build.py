"""Builds gpupq_amd/libgpupq.so from gpupq_amd/csrc/*.hip with hipcc for gfx950.

    python gpupq_amd/build.py        (or __graft_entry__.build(); `python -m gpupq_amd.build`
                                      would first import the package, which needs the library)

Every translation unit is compiled with -ffp-contract=off (bit-identical float
arithmetic with the CPU reference: no fused multiply-adds) and with packed fp32
instructions disabled.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgpupq.so")
OBJ_DIR = os.path.join(HERE, "_objs")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GPUPQ_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), res.stderr))
    noise = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    return obj, "\n".join(noise)


def build(verbose=True):
    os.makedirs(OBJ_DIR, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not sources:
        raise RuntimeError("no sources under %s" % CSRC)
    jobs = min(len(sources), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, sources))
    for obj, msg in results:
        if msg and verbose:
            print(msg, file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [o for o, _ in results]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


def build_model(src, out, verbose=True):
    """Compiles a caller's model kernels (`src`, HIP) against the public headers under include/
    into the shared library `out`.  Nothing of libgpupq is linked: the model reaches the queue
    through gpq_queue_device_view at run time, and its kernels inline the device verbs of
    include/gpupq_device.h."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I" + inc, "-shared", src, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), res.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build()
