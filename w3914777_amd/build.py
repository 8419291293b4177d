"""Build libamps.so (the C-ABI engine) for gfx950 with hipcc, in-tree.

Every .hip/.cpp file under csrc/ is compiled to an object under _build/ and
linked into w3914777_amd/libamps.so.  Packed FP32 instructions are disabled
for all device code (-packed-fp32-ops).  The command-line drivers
(drivers/amps_run.cpp, host code over the C ABI) are linked against it into
w3914777_amd/bin/amps_run.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libamps.so")
DRIVER_SRC = os.path.join(HERE, "drivers", "amps_run.cpp")
DRIVER = os.path.join(HERE, "bin", "amps_run")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include")]
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE if src.endswith(".hip") else []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, r.stdout + r.stderr, obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for src, rc, out, obj in ex.map(_compile, srcs):
            warn = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if rc != 0:
                sys.stderr.write(out)
                raise RuntimeError("hipcc failed on " + src)
            if verbose and warn:
                sys.stderr.write("\n".join(warn) + "\n")
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    build_driver()
    build_probes()
    return LIB


def build_probes():
    """The VALU issue-cost probe (probes/valu_rates.hip), a measurement tool
    beside the library."""
    src = os.path.join(HERE, "probes", "valu_rates.hip")
    out = os.path.join(HERE, "probes", "libvalu_rates.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-shared"] + DEVICE + [src, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("probe build failed")
    os.replace(tmp, out)
    return out


def build_driver():
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    tmp = DRIVER + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(os.path.dirname(HERE), "include"), DRIVER_SRC,
           "-L" + HERE, "-lamps", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("driver build failed")
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
