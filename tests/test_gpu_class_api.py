"""The reference's unit tests restated against this build's C++ class API
(tests/cpp/class_api_test.cpp, linked to libvpd.so; every computation on the GPU)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "vumat-pd_amd", "build", "class_api_test")


def test_class_api_restated_reference_tests():
    assert os.path.exists(BIN), "build the C++ class-API test: make -C vumat-pd_amd"
    print("[gpu-case] C++ class API", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "class API tests passed" in r.stdout
