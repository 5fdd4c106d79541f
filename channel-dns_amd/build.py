"""Build the gfx950 backend library in place: channel-dns_amd/libcdns_hip.so."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["api.hip", "transforms.hip", "transforms_rf.hip", "transforms_y.hip", "spectral.hip", "solver.hip", "adjoint.hip", "diagnostics.hip",
           "transforms_dct.hip"]
OUT = os.path.join(HERE, "libcdns_hip.so")
OBJDIR = os.path.join(HERE, "_build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core accumulators in ordinary VGPRs: with AGPR accumulators the y kernels copied every
    # accumulator out and back in each k-step
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode, r.stdout + r.stderr


def build(verbose=False, jobs=None):
    """Compile every source to an object (in parallel), then link the shared library."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    objs, cmds = [], []
    for s in SOURCES:
        o = os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        cmds.append([hipcc] + HIPCC_FLAGS + ["-c", "-o", o, os.path.join(HERE, "csrc", s)])
    with ThreadPoolExecutor(max_workers=jobs or min(len(cmds), os.cpu_count() or 1)) as ex:
        results = list(ex.map(_run, cmds))
    for (rc, log), s in zip(results, SOURCES):
        if rc != 0:
            sys.stderr.write(log)
            raise RuntimeError(f"hipcc failed on {s}")
        if verbose:
            sys.stderr.write(log)
    tmp = OUT + ".tmp"
    rc, log = _run([hipcc] + HIPCC_FLAGS + ["-shared", "-o", tmp] + objs)
    if rc != 0:
        sys.stderr.write(log)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, OUT)
    # the C++ mirror of the reference's Simulation loop over the C ABI (host/run_steps)
    exe = os.path.join(HERE, "host", "run_steps")
    rc, log = _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(os.path.dirname(HERE), "include"), os.path.join(HERE, "host", "run_steps.cpp"),
                    "-L", HERE, "-lcdns_hip", "-Wl,-rpath,$ORIGIN/..", "-o", exe])
    if rc != 0:
        sys.stderr.write(log)
        raise RuntimeError("g++ failed on host/run_steps.cpp")
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
