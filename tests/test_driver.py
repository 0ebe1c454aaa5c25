"""The drop-in driver `coupled_cluster_doubles_amd/ccd`: the reference's
options (src/parse.cpp:131-179) and report lines (src/main.cpp:482-700).

CPU tests cover option handling that ends before a device is touched; GPU
tests run whole solves through the binary and hold the printed energies to
the oracle.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle.pyoracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CCD = os.path.join(ROOT, "coupled_cluster_doubles_amd", "ccd")


def run(*args, timeout=120):
    return subprocess.run([CCD, *args], capture_output=True, text=True, timeout=timeout)


def report(out):
    """The key: value lines of the report (last value of each key)."""
    vals = {}
    for line in out.splitlines():
        m = re.match(r"^([A-Za-z_0-9/ ]+): (\S+)$", line)
        if m:
            vals[m.group(1)] = m.group(2)
    return vals


def iteration_energies(out):
    return np.array([float(x) for x in re.findall(r"^corrCCD: (\S+)$", out, re.M)])


# ---------------------------------------------------------------- CPU

def test_help_lists_the_reference_options():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--config", "--saveMemory", "--pairing", "--infMatter", "--electronGas", "--qDots",
                "--density", "--nParticleShells", "--nParticles", "--nPairStates", "--r_s", "--nShells",
                "--tolerance", "--tzMax", "--xi", "--precision"):
        assert opt in r.stdout, opt


def test_quantum_dots_are_not_ready_like_the_reference():
    r = run("--qDots")
    assert r.returncode == 1 and "not ready" in r.stdout


def test_missing_config_file_is_an_error():
    r = run("-c", "/nonexistent/ccd.cfg")
    assert r.returncode == 1 and "cannot open config file" in r.stderr


def test_bad_precision_is_rejected():
    r = run("--precision", "fp16")
    assert r.returncode != 0 and "unknown precision" in r.stderr


def test_symmetric_nuclear_matter_is_rejected_before_the_device():
    r = run("--infMatter", "-z", "2")
    assert r.returncode == 1 and "pure neutron matter" in r.stderr


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_driver_pairing_matches_oracle(require_gpu):
    r = run("--pairing", "-l", "8", "-n", "8", "-g", "0.5", "-x", "1.0", "-t", "1e-12", "-v", "1")
    assert r.returncode == 0, r.stderr
    v = report(r.stdout)
    o = Oracle.pairing(8, 8, 0.5, 1.0)
    e_o, stop, _ = o.solve(1e-12)
    assert float(v["Reference Energy"]) == 11.0
    assert abs(float(v["MBPT2Corr"]) - o.mbpt2()) <= 1e-13
    assert int(v["iterations"]) == len(e_o)
    assert abs(float(v["CCDCorr"]) - e_o[-1]) <= 1e-10 * abs(e_o[-1])
    e = iteration_energies(r.stdout)
    assert len(e) == len(e_o) and np.max(np.abs(e - e_o) / np.abs(e_o)) < 1e-10
    assert float(v["E_CCD"]) == pytest.approx(11.0 + e_o[-1], rel=1e-12)


@pytest.mark.gpu
def test_driver_electron_gas_stored_and_matrix_free(require_gpu, tmp_path):
    """-m 1 prints the same energies as -m 0; a config file supplies the
    basis and the command line overrides its tolerance (src/parse.cpp:209-229)."""
    cfg = tmp_path / "heg.cfg"
    cfg.write_text("--electronGas -r 1.0 -s 19 -p 2 -t 1e-3\n")
    r0 = run("-c", str(cfg), "-t", "1e-10", "-v", "1")
    r1 = run("-c", str(cfg), "-t", "1e-10", "-v", "1", "-m", "1")
    assert r0.returncode == 0 and r1.returncode == 0, r0.stderr + r1.stderr
    e0, e1 = iteration_energies(r0.stdout), iteration_energies(r1.stdout)
    assert len(e0) > 5 and np.array_equal(e0, e1)
    o = Oracle.electron_gas(1.0, 19, 2)
    e_o, _, _ = o.solve(1e-10)
    v = report(r0.stdout)
    assert abs(int(v["iterations"]) - len(e_o)) <= 1
    assert abs(float(v["CCDCorr"]) - e_o[-1]) <= 1e-10 * abs(e_o[-1])
    assert abs(float(v["Reference Energy"]) - o.eref()) <= 1e-14 * abs(o.eref())


@pytest.mark.gpu
@pytest.mark.parametrize("save_memory", ["0", "1"])
def test_driver_inf_matter_published_energy(require_gpu, save_memory):
    r = run("--infMatter", "-d", "0.08", "-z", "1", "-s", "4", "-p", "2", "-t", "1e-8", "-m", save_memory)
    assert r.returncode == 0, r.stderr
    v = report(r.stdout)
    assert abs(float(v["E_CCD/A"]) - 10.155078) < 5e-7


@pytest.mark.gpu
def test_driver_fp32_contractions_stay_close(require_gpu):
    r64 = run("--electronGas", "-r", "1.0", "-s", "19", "-p", "2", "-t", "1e-8")
    r32 = run("--electronGas", "-r", "1.0", "-s", "19", "-p", "2", "-t", "1e-8", "--precision", "fp32")
    assert r64.returncode == 0 and r32.returncode == 0
    a, b = float(report(r64.stdout)["CCDCorr/A"]), float(report(r32.stdout)["CCDCorr/A"])
    assert abs(a - b) <= 1e-6
