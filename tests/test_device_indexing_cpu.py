"""CPU check of the device-side index arithmetic before it runs on the GPU.

Builds tests/cpu_checks/check_indexing.cpp (host-only, includes the kernels'
u1_device.h) and runs it: magic division against '/', and every staple
neighbour position of every site of several geometries (2^4 ... 48^4) against
brute-force coordinates and the half-plane bounds.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_device_indexing_on_cpu(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "check_indexing"
    subprocess.run([hipcc, "-O2", "-std=c++17", os.path.join(HERE, "cpu_checks", "check_indexing.cpp"), "-o",
                    str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "indexing ok" in out.stdout


def test_fma_proposal_matches_two_step_form(tmp_path):
    """The kernels' one-FMA Metropolis proposal offset is bit-identical to the
    oracle's two-step a (2u - 1), and the binary32 kernels' binary64 wrap
    wrap_near_d lands in [-pi, pi] on the angle wrap_d gives, within a few
    binary64 ulps (tests/cpu_checks/check_proposal.cpp)."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "check_proposal"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-mfma", "-I",
                    os.path.join(os.path.dirname(HERE), "latticeu1_amd", "csrc"),
                    os.path.join(HERE, "cpu_checks", "check_proposal.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "proposal ok" in out.stdout and "wrap ok" in out.stdout


def test_fused_pass_schedule_and_dependencies(tmp_path):
    """The fused parity pass (u1_chain.h): every (phase, plane, chunk) once in
    the per-XCD lists, every dependency strictly earlier in schedule time, and
    an emulation of the launch under random interleavings (reads of unwritten
    angles at start, the waits of chain_deps, reads of the new angles, writes)
    that reproduces four sequential phases exactly and catches a dropped
    dependency (tests/cpu_checks/check_chain.cpp)."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "check_chain"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", os.path.join(os.path.dirname(HERE), "latticeu1_amd", "csrc"),
                    os.path.join(HERE, "cpu_checks", "check_chain.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "schedule ok" in out.stdout and "emulation ok" in out.stdout
