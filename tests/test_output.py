"""Text dumps (DumpParticles / DumpFields, Write.f90) byte for byte against
the reference's own writers compiled with amdflang (golden fixture written by
tests/golden/make_golden.py), on values spread over many decades, with zeros,
negative zeros, rounding ties and three-digit exponents."""
import os

import numpy as np
import pytest

from picard_amd import output

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_dump.npz")


def _z():
    if not os.path.exists(GOLDEN):
        pytest.skip("golden fixture ref_dump.npz missing")
    return np.load(GOLDEN, allow_pickle=False)


def test_e10_4_known_answers():
    # amdflang's E10.4 of these values (ties to even on the exact binary value)
    cases = [(0.015625, "0.1562E-01"), (-0.015625, "-.1562E-01"), (0.5, "0.5000E+00"),
             (9.99995, "0.1000E+02"), (1e100, "0.1000+101"), (1e-99, "0.1000E-98"),
             (5e-324, "0.4941-323"), (0.0, "0.0000E+00"), (-0.0, "-.0000E+00"),
             (0.12345, "0.1235E+00"), (1.7976931348623157e308, "0.1798+309")]
    got = output.format_e10_4([v for v, _ in cases])
    assert got.tolist() == [t for _, t in cases]


def test_particles_dump_matches_reference(tmp_path):
    z = _z()
    name = output.write_particles(str(tmp_path), z["rp"], z["sp"], 7, 0)
    assert os.path.basename(name) == "particles_00007_00000"
    with open(name, "rb") as f:
        assert f.read() == z["particles_txt"].tobytes()


def test_fields_dump_matches_reference(tmp_path):
    z = _z()
    names = output.write_fields(str(tmp_path), z["U"], z["N"], z["F"], z["E"], 2, 0)
    got = {os.path.basename(n) for n in names}
    assert got == set(z["field_names"].tolist())
    for n in names:
        with open(n, "rb") as f:
            assert f.read() == z["txt_" + os.path.basename(n)].tobytes(), n


OUTPUT_CHECK = os.path.join(os.path.dirname(__file__), "fortran", "output_check")


def test_fortran_host_dumps_match_reference(tmp_path):
    """The Fortran host's writers (picard_amd/fortran/picard_output.f90) give
    the reference's files byte for byte."""
    import subprocess
    z = _z()
    if not os.path.exists(OUTPUT_CHECK):
        pytest.skip("tests/fortran/output_check is not built")
    nx, ny, nz, S = [int(v) for v in z["dims"]]
    rp, sp = z["rp"], z["sp"]
    with open(tmp_path / "dims.txt", "w") as f:
        f.write(f"{rp.shape[0]} {nx} {ny} {nz} {S} 7 2\n")
    np.ascontiguousarray(rp, np.float64).tofile(tmp_path / "rp.bin")
    np.ascontiguousarray(sp, np.int32).tofile(tmp_path / "sp.bin")
    for k in ("U", "E", "N", "F"):
        np.ascontiguousarray(z[k], np.float64).tofile(tmp_path / f"{k}.bin")
    subprocess.run([OUTPUT_CHECK, str(tmp_path)], check=True, timeout=60)
    with open(tmp_path / "save" / "particles" / "particles_00007_00000", "rb") as f:
        assert f.read() == z["particles_txt"].tobytes()
    for name in z["field_names"].tolist():
        with open(tmp_path / "save" / "fields" / name, "rb") as f:
            assert f.read() == z["txt_" + name].tobytes(), name
