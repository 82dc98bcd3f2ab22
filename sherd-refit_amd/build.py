"""Builds libsherd_hip.so for gfx950 from csrc/ (hipcc, one object per translation unit).

    python sherd-refit_amd/build.py

The library lands in sherd-refit_amd/sherd_hip/lib/ (in-tree, git-ignored); the counter-calibration
program in sherd-refit_amd/calib/copy_probe (git-ignored).
"""

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "sherd_hip", "lib")
LIB = os.path.join(LIB_DIR, "libsherd_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no fused multiply-adds, so the device performs the reference's operations one
# at a time. Packed FP32 is disabled for all HIP code. f32 division and square root stay correctly
# rounded (the HIP default, spelled out).
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def sources():
    return sorted(f for f in os.listdir(SRC) if f.endswith(".hip") or f.endswith(".cpp"))


def _compile(name):
    out = os.path.join(OBJ, os.path.splitext(name)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, name), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"  # hipcc's host pass does not know the feature; harmless
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    return name, r.returncode, msg, out


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for name, rc, msg, out in ex.map(_compile, sources()):
            if rc != 0:
                raise RuntimeError(f"hipcc failed on {name}:\n{msg}")
            if verbose and msg.strip():
                print(msg)
            objs.append(out)
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    # The counter-calibration program (not part of the library): a known-size copy kernel.
    probe = os.path.join(HERE, "calib", "copy_probe")
    r = subprocess.run([HIPCC] + FLAGS + [os.path.join(HERE, "calib", "copy_probe.hip"), "-o", probe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on calib/copy_probe.hip:\n" + r.stdout + r.stderr)
    return LIB


if __name__ == "__main__":
    print(build())
    sys.exit(0)
