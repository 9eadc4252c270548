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


PARAMS = os.path.join(os.path.dirname(__file__), "golden", "ref_parameters.npz")


def _parameter_cases():
    import json
    from picard_amd import host
    if not os.path.exists(PARAMS):
        pytest.skip("golden fixture ref_parameters.npz missing")
    z = np.load(PARAMS, allow_pickle=False)
    for t in range(int(z["ncase"])):
        desc = json.loads(str(z[f"config{t}"]))
        cfg = host.Config(**{k: (tuple(v) if isinstance(v, list) else v) for k, v in desc.items()})
        yield t, cfg, tuple(int(v) for v in z[f"procs{t}"]), z[f"parameters{t}"].tobytes()


def test_parameters_dat_matches_reference(tmp_path):
    """save/parameters.dat (WriteWarnings.f90:130-256) byte for byte against
    the reference's routine compiled with amdflang: the 32^3 case with
    picard.f90's Haser parameters, the 256^3 workload, and a case with an I4
    overflow, three-digit exponents, negative values and mixed flags."""
    from picard_amd import host
    n = 0
    for t, cfg, procs, want in _parameter_cases():
        d = host.derive(cfg)
        name = output.write_parameters(str(tmp_path / f"c{t}"), cfg, d, procs)
        with open(name, "rb") as f:
            assert f.read() == want, f"case {t}"
        n += 1
    assert n == 3


def test_e12_6_known_answers():
    cases = [(0.625e9, "0.625000E+09"), (-1.60217646e-19, "-.160218E-18"), (0.0, "0.000000E+00"),
             (1e-300, "0.100000-299"), (9.9999995e5, "0.100000E+07"), (7.927266729303253e-07,
                                                                      "0.792727E-06")]
    assert output.format_e([v for v, _ in cases], 6).tolist() == [t for _, t in cases]


def test_fortran_host_parameters_match_reference(tmp_path):
    """The Fortran host's parameters.dat writer (dump_parameters in
    picard_amd/fortran/picard_output.f90) gives the reference's bytes."""
    import subprocess
    from picard_amd import host
    if not os.path.exists(OUTPUT_CHECK):
        pytest.skip("tests/fortran/output_check is not built")
    n = 0
    for t, cfg, procs, want in _parameter_cases():
        d = host.derive(cfg)
        w = tmp_path / f"c{t}"
        w.mkdir()
        reals = np.concatenate([np.asarray(v, np.float64)[:8] for v in
                                (d.n2p, d.mass, d.charge, d.density, d.kelvin, cfg.Qn, cfg.vn,
                                 cfg.Ek, cfg.nu_i, cfg.nu_d)] +
                               [np.asarray(d.ve, np.float64)[:4], np.asarray(d.B0, np.float64)[:4],
                                np.asarray(d.dxyz, np.float64)[:3],
                                np.array([cfg.gmin[0], cfg.gmax[0], cfg.gmin[1], cfg.gmax[1],
                                          cfg.gmin[2], cfg.gmax[2], d.dt])])
        ints = np.concatenate([np.asarray(d.ppc, np.int32)[:8],
                               np.array([1 if g else 0 for g in cfg.gyration], np.int32),
                               np.array([cfg.nx, cfg.ny, cfg.nz], np.int32),
                               np.array(procs, np.int32),
                               np.array([1 if f else 0 for f in cfg.flow], np.int32)])
        with open(w / "params.bin", "wb") as f:
            f.write(reals.tobytes())
            f.write(ints.tobytes())
        subprocess.run([OUTPUT_CHECK, str(w), "params"], check=True, timeout=60)
        with open(w / "save" / "parameters.dat", "rb") as f:
            assert f.read() == want, f"case {t}"
        n += 1
    assert n == 3
