"""The C++ header shim compiles with a plain host compiler (no HIP, no
Kokkos headers) and the built driver links against libnukexc_hip.so."""
import os
import subprocess

from conftest import ROOT


def test_shim_headers_compile_without_hip():
    src = os.path.join(ROOT, "tests", "cpp", "shim_driver.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_shim_driver_links():
    exe = os.path.join(ROOT, "tests", "cpp", "shim_driver")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    r = subprocess.run(["ldd", exe], capture_output=True, text=True)
    assert "libnukexc_hip.so" in r.stdout and "not found" not in r.stdout
