"""Build the tracker's native libraries in-tree.

librecon_amd.so   HIP kernels + C ABI (hipcc, gfx950), including the seeded
                  synthetic RGB-D renderer (host and device)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["rt_kernels.hip", "rt_tracker.hip", "rt_batch.hip", "cnn.hip", "synth.hip", "rt_sensor.hip", "rt_trajectory.hip", "rt_tsdf.hip", "rt_sequence.hip"]
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    # every FP32 value is the rounded result of the written operation sequence
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd[:3]))
    return out


def build_hip():
    """Compile each source to an object in parallel, then link the library."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]

    def obj(src):
        o = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run([hipcc] + compile_flags + ["-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", o])
        return o

    with ThreadPoolExecutor(max_workers=min(len(HIP_SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(obj, HIP_SOURCES))
    out = os.path.join(HERE, "librecon_amd.so")
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", out] + objs)
    return out


def build_all():
    return [build_hip()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
