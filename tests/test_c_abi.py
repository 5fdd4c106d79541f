"""The ABI as a C caller sees it: tests/c_abi/abi_check.c includes include/cdns.h, takes every
entry point into a table of exactly typed function pointers, is compiled as strict C99 with
warnings as errors and linked against the built library, then checks what needs no device
(null-argument handling, the host-only engine's known answer)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "channel-dns_amd")


def _declared_entry_points():
    with open(os.path.join(ROOT, "include", "cdns.h")) as f:
        text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
    return sorted(set(re.findall(r"\b(cdns_[a-z0-9_]+)\s*\(", text)))


def test_c_caller_compiles_links_and_runs(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libcdns_hip.so")):
        pytest.skip("library not built")
    exe = str(tmp_path / "abi_check")
    src = os.path.join(ROOT, "tests", "c_abi", "abi_check.c")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
                    src, "-L", LIBDIR, "-lcdns_hip", f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(src) as f:
        table = f.read()
    missing = [n for n in _declared_entry_points() if not re.search(rf"\b{n}\b", table.split("cdns_api api")[1])]
    assert not missing, f"entry points of cdns.h not in the typed table: {missing}"


def test_cpp_simulation_binding_is_built():
    """channel-dns_amd/host/run_steps (the C++ Simulation loop over the ABI, built by build.py with
    -Wall -Wextra -Werror) exists and rejects a wrong command line before touching a device."""
    exe = os.path.join(LIBDIR, "host", "run_steps")
    if not os.path.exists(exe):
        pytest.skip("not built")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
