"""bench.py's host-side bookkeeping (no GPU work): which device kernel a launch tuning runs,
and the committed HBM traffic figure it reports for it."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def tuning(pass_, window, form):
    return {pass_ + "_window": window, pass_ + "_window_plan": form}


def test_bond_pass_device_kernels():
    # The names rocprofv3 gives the step's bond-pass kernels in each gather form.
    f = lambda w, m: bench.bond_pass_device_kernel("defgrad_bond_pass", tuning("defgrad", w, m))
    g = lambda w, m: bench.bond_pass_device_kernel("force_bond_pass", tuning("force", w, m))
    assert f(0, 3) == "k_defgrad<true>"
    assert f(80, 0) == "k_defgrad_w<true, 80>"
    assert f(128, 1) == "k_defgrad_p<true, 128, false>"
    assert f(96, 2) == "k_defgrad_p<true, 96, true>"
    assert f(480, 3) == "k_defgrad_g<true, 480>"
    assert g(0, 1) == "k_force<true, true, true, false>"
    assert g(80, 0) == "k_force_w<true, true, true, 80>"
    assert g(64, 1) == "k_force_p<true, true, true, 64, false>"
    assert g(512, 3) == "k_force_g<true, true, true, 512>"
    assert f(64, 4) == "k_defgrad_m<true, 64>"
    assert g(16, 4) == "k_force_m<true, true, true, 16>"


def test_traffic_only_for_the_profiled_kernel():
    # A committed traffic summary is reported only while the tuning runs the kernel it was
    # measured on.
    for kernel in ("defgrad_bond_pass", "force_bond_pass"):
        path = os.path.join(ROOT, "profiles", "traffic_%s_jc_tensile_bar_500k.json" % kernel)
        with open(path) as fh:
            data = json.load(fh)
        got = bench.measured_traffic(kernel, "jc_tensile_bar_500k", data["device_kernel"])
        assert got == data["hbm_bytes_per_launch"] > 0
        assert bench.measured_traffic(kernel, "jc_tensile_bar_500k", "k_other<true>") is None
        assert bench.measured_traffic(kernel, "another_workload", data["device_kernel"]) is None
