"""The Fortran drop-in layer on the GPU: the reference's subroutine names and
argument lists, called from Fortran, bit-identical to the oracle; and the
Fortran host program running the main loop."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def _run(cmd, timeout, cwd=None):
    print(f"[gpu-fortran] running {' '.join(cmd)}", flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd=cwd)
    print(out.stdout[-3000:], out.stderr[-2000:], flush=True)
    return out


def test_fortran_dropin_matches_oracle():
    exe = os.path.join(ROOT, "tests", "fortran", "dropin_check")
    assert os.path.exists(exe), "build() compiles tests/fortran/dropin_check"
    out = _run([exe, "4"], 240)
    assert out.returncode == 0
    assert "DROPIN OK" in out.stdout


@pytest.mark.parametrize("case,arg", [("xflow", "xflow/yflow/zflow"), ("dxyz", "dxyz"),
                                      ("tB", "tB"), ("max_per_proc", "max_per_proc"),
                                      ("n2p", "n2p"), ("E0", "E0"),
                                      ("myid", "iprocs/jprocs/kprocs/myid")])
def test_fortran_dropin_rejects_changed_arguments(case, arg):
    """A drop-in passed an argument that differs from the domain registered
    at picard_dropin_init (one flag, one cell size, one bit of tB, ...)
    stops with an error naming it, before any device work, instead of
    silently using the registered value."""
    exe = os.path.join(ROOT, "tests", "fortran", "dropin_check")
    out = _run([exe, "mismatch", case], 120)
    assert out.returncode == 2, out.returncode
    assert f"argument {arg}" in out.stdout and "differs from the domain" in out.stdout
    assert "MISMATCH ACCEPTED" not in out.stdout


def test_fortran_host_runs(tmp_path):
    exe = os.path.join(ROOT, "picard_amd", "picard_host")
    assert os.path.exists(exe), "build() compiles picard_amd/picard_host"
    nml = tmp_path / "run.nml"
    nml.write_text("&picard_run\n nx=16, ny=16, nz=16, steps=3, xmin=-80, xmax=80, ymin=-80, "
                   "ymax=80, zmin=-80, zmax=80 /\n")
    out = _run([exe, str(nml)], 240, cwd=str(tmp_path))
    assert out.returncode == 0
    lines = [ln for ln in out.stdout.splitlines() if ln.startswith("step")]
    assert len(lines) == 3
    assert "particle-steps/s" in out.stdout
    # save/parameters.dat as the reference writes it at setup: the Python
    # writer (pinned byte for byte against the reference's WriteWarnings) on
    # the same run, with picard.f90's Haser parameters
    from picard_amd import host, output
    ec = host.EC
    cfg = host.Config(nx=16, ny=16, nz=16, gmin=(-80.0,) * 3, gmax=(80.0,) * 3,
                      Qn=(0.0, 0.0, 400.0 ** (-3) * 1.99e26, 0, 0, 0, 0, 0),
                      vn=(0.0, 0.0, 629.0, 0, 0, 0, 0, 0), Ek=(0.0, 0.0, 1.52e1 * ec, 0, 0, 0, 0, 0),
                      nu_i=(0.0, 0.0, 1.26e-7, 0, 0, 0, 0, 0), nu_d=(0.0, 0.0, 1.26e-7, 0, 0, 0, 0, 0))
    with open(tmp_path / "save" / "parameters.dat") as f:
        assert f.read() == output.parameters_text(cfg, host.derive(cfg))
