"""CPU checks of the C ABI: libfhd.so loads, exports every function include/fhd.h declares, the
ctypes struct matches the compiled one, and the data.main front end (pure host code) parses
fluam's grammar (src/loadDataMain.cu:173-683) and rejects what fluam rejects."""
import ctypes as C
import os
import re
import subprocess

import pytest

import fluam_amd as F
from fluam_amd._abi import FhdParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DM = os.path.join(ROOT, "tests", "golden", "datamain")


def _declared_functions():
    hdr = open(os.path.join(ROOT, "include", "fhd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(fhd_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    L = F.lib()
    names = _declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}$", out, flags=re.M), n


def test_params_struct_size_matches():
    L = F.lib()
    L.fhd_params_size.restype = C.c_ulong
    assert L.fhd_params_size() == C.sizeof(FhdParams)


def test_parse_incompressible_datamain():
    p = F.load_datamain(os.path.join(DM, "incompressible_32.data.main"))
    assert p.scheme == F.SCHEME_INCOMPRESSIBLE
    assert (p.mx, p.my, p.mz) == (32, 32, 32)
    assert p.temperature == pytest.approx(1e-3)
    assert p.dt == pytest.approx(0.05)
    assert p.numsteps == 100 and p.samplefreq == 50 and p.seed == 1
    assert p.np == 0
    assert p.outputname == b"out/incompressible"


def test_parse_stokes_datamain_sets_particles_and_extensions():
    p = F.load_datamain(os.path.join(DM, "stokes_128.data.main"))
    assert p.scheme == F.SCHEME_STOKES_LIMIT
    assert p.np == 65536
    assert p.kernel_points == 4
    assert p.wca_sigma == pytest.approx(1.0)
    # epsilon defaults to kT (src/initForcesNonBonded.cu:25)
    assert p.wca_epsilon == pytest.approx(1e-3)
    assert p.cutoff == pytest.approx(2 ** (1 / 6))
    assert p.initial_positions == 0


def test_parse_qnb_volume_default():
    p = F.load_datamain(os.path.join(DM, "qnb_64.data.main"))
    assert p.scheme == F.SCHEME_QUASI_NEUTRALLY_BUOYANT
    assert p.volume_particle == pytest.approx(8.0)  # src/createParticles.cpp:141-143
    assert p.mass == 0.0


def test_unknown_keyword_is_an_error():
    with pytest.raises(F.FhdError, match="ERROR INPUT FILE"):
        F.load_datamain(os.path.join(DM, "bad_keyword.data.main"))


def test_out_of_scope_scheme_is_reported(tmp_path):
    f = tmp_path / "d.main"
    f.write_text("binaryMixture\ncells 8 8 8\n")
    with pytest.raises(F.FhdError, match="outside this library"):
        F.load_datamain(str(f))


def test_samplefreq_zero_fails_cleanly(tmp_path):
    # fluam divides by samplefreq (src/runSchemeIncompressible.cu:141); here it is a clean error
    f = tmp_path / "d.main"
    f.write_text("incompressible\ncells 8 8 8\ncelldimension 8 8 8\ndt 0.1\nnumsteps 2\n")
    F.load_datamain(str(f))
    assert F.scheme_incompressible() is False
    assert b"samplefreq" in F.lib().fhd_last_error()


def test_comment_lines_are_skipped(tmp_path):
    f = tmp_path / "d.main"
    f.write_text("# a comment incompressible cells 1 1 1\nstokesLimit\n#x\ncells 16 16 16\n")
    p = F.load_datamain(str(f))
    assert p.scheme == F.SCHEME_STOKES_LIMIT and p.mx == 16


def _write_reference_layout(path, cells, box, v, dens):
    """The layout of src/saveFluidFinalConfiguration.cu:36-71, written independently of the
    package's writer (precision 15, as fluam writes it)."""
    mx, my, mz = cells
    N = mx * my * mz
    lines = [f"Number of cells {N}", "mx my mz", f"{mx} {my} {mz}", "lx ly lz", f"{box[0]} {box[1]} {box[2]}",
             f"Density {dens}", "Pressure parameters 0 0 0", "Thermostat 0", "Temperature 0.001",
             "Cells properties ", "Density ", "Velocity"]
    for i in range(N):
        lines.append(f"{dens}")
        lines.append(f"{v[i]:.15g} {v[N + i]:.15g} {v[2 * N + i]:.15g}")
    path.write_text("\n".join(lines) + "\n")


def test_read_fluid_file_reference_layout(tmp_path):
    import numpy as np
    rng = np.random.default_rng(3)
    cells = (4, 6, 5)
    v = rng.standard_normal(3 * 120)
    f = tmp_path / "run.fluidFinalConfiguration"
    _write_reference_layout(f, cells, (4.0, 6.0, 5.0), v, 1.25)
    got, dens = F.read_fluid_file(f, cells)
    assert dens == 1.25
    assert np.allclose(got, v, rtol=1e-14, atol=0)
    # the package's writer round-trips through the same reader exactly
    g = tmp_path / "w.fluidFinalConfiguration"
    F.write_fluid_file(g, cells, (4.0, 6.0, 5.0), v, densfluid=0.75)
    got2, dens2 = F.read_fluid_file(g, cells)
    assert dens2 == 0.75 and np.array_equal(got2, v)


def test_read_fluid_file_rejects_wrong_grid_and_truncation(tmp_path):
    import numpy as np
    v = np.zeros(3 * 64)
    f = tmp_path / "a.fluidFinalConfiguration"
    _write_reference_layout(f, (4, 4, 4), (4.0, 4.0, 4.0), v, 1.0)
    with pytest.raises(F.FhdError, match="grid differs"):
        F.read_fluid_file(f, (4, 4, 5))
    text = f.read_text().splitlines()
    (tmp_path / "b").write_text("\n".join(text[:-10]) + "\n")
    with pytest.raises(F.FhdError, match="truncated"):
        F.read_fluid_file(tmp_path / "b", (4, 4, 4))


def test_parse_fluid_restart_keywords(tmp_path):
    f = tmp_path / "d.main"
    f.write_text("incompressible\ncells 8 8 8\ncelldimension 8 8 8\nfluid saved.fluidFinalConfiguration\n"
                 "initialStep 42\n")
    p = F.load_datamain(str(f))
    assert p.fluid == b"saved.fluidFinalConfiguration" and p.initial_step == 42


def test_parse_bonded_forces_keywords(tmp_path):
    f = tmp_path / "d.main"
    f.write_text("stokesLimit\ncells 8 8 8\ncelldimension 8 8 8\nnumberparticles 2\n"
                 "bondedForces springs.dat\nbondedForcesVersion 0\n")
    p = F.load_datamain(str(f))
    assert p.bonded_forces == b"springs.dat"


def test_old_bond_file_layout_is_reported(tmp_path):
    f = tmp_path / "d.main"
    f.write_text("stokesLimit\ncells 8 8 8\nbondedForcesVersion 1\n")
    with pytest.raises(F.FhdError, match="bondedForcesVersion"):
        F.load_datamain(str(f))


def test_parse_structure_factor_keyword(tmp_path):
    f = tmp_path / "d.main"
    f.write_text("incompressible\ncells 8 8 8\nstructureFactor 1\n")
    assert F.load_datamain(str(f)).structure_factor == 1
