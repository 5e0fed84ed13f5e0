"""The C++ header shim (include/nukexc/*.hpp) driven like the reference's own
unit tests, from a plain g++ binary linked to libnukexc_hip.so."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_shim_driver():
    exe = os.path.join(ROOT, "tests", "cpp", "shim_driver")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "shim OK" in r.stdout
