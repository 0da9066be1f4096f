"""fluam's input semantics on the CPU side: the example data.main files fluam ships
(bin/inputfiles/data.main.{incompressible,quasiNeutrallyBuoyant,stokesLimit}, kept here as
fixtures), loadparticles with a coordinates file whose first number is the particle count
(src/createParticles.cpp:81-100), the optional velocities file (src/createParticles.cpp:103-120)
and the initfluid options (src/initializeFluid.cpp:31-160), checked through the library's parser
and the oracle."""
import os

import numpy as np
import pytest

import fluam_amd as F
import oracle_lib as O
from fluam_amd._abi import (make_params, SCHEME_INCOMPRESSIBLE, SCHEME_QUASI_NEUTRALLY_BUOYANT,
                            SCHEME_STOKES_LIMIT)

EXAMPLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "datamain", "fluam_examples")


def _have_lib():
    return os.path.exists(F.LIB_PATH)


@pytest.mark.skipif(not _have_lib(), reason="libfhd.so not built")
@pytest.mark.parametrize("name,scheme,np_,extra", [
    ("incompressible", SCHEME_INCOMPRESSIBLE, 0, {}),
    ("quasiNeutrallyBuoyant", SCHEME_QUASI_NEUTRALLY_BUOYANT, 1,
     {"mass": 8000.0, "max_part_in_cell": 20, "max_part_in_cell_nonbonded": 20}),
    ("stokesLimit", SCHEME_STOKES_LIMIT, 1, {"max_part_in_cell": 20, "max_part_in_cell_nonbonded": 20}),
])
def test_fluam_example_inputs_parse(name, scheme, np_, extra):
    p = F.load_datamain(os.path.join(EXAMPLES, "data.main." + name))
    assert p.scheme == scheme
    assert (p.mx, p.my, p.mz) == (32, 32, 32)
    assert (p.lx, p.ly, p.lz) == (320.0, 320.0, 320.0)
    assert p.densfluid == 1.0 and p.shearviscosity == 1.0 and p.temperature == 0.1
    assert p.dt == 0.25 and p.numsteps == 1000 and p.samplefreq == 50
    assert p.initfluid == 1
    assert p.np == np_
    if np_:
        assert abs(p.cutoff - 22.4492409662) < 1e-12
    for k, v in extra.items():
        assert getattr(p, k) == v, k
    assert p.outputname == b"../data/run1"


def _write(path, rows):
    with open(path, "w") as f:
        f.write(f"{len(rows)}\n")
        for r in rows:
            f.write(" ".join(repr(float(x)) for x in r) + "\n")


@pytest.mark.skipif(not _have_lib(), reason="libfhd.so not built")
def test_coordinates_file_sets_the_particle_count(tmp_path):
    coords = tmp_path / "p.coor"
    vels = tmp_path / "p.vel"
    q = [(1.0, -2.0, 3.5), (-4.25, 0.5, 0.0), (2.0, 2.0, -6.0)]
    u = [(0.1, 0.2, 0.3), (-0.4, 0.5, -0.6), (0.7, -0.8, 0.9)]
    _write(coords, q)
    _write(vels, u)
    dm = tmp_path / "data.main"
    dm.write_text("stokesLimit\nnumberparticles 7\ncells 16 16 16\ncelldimension 16 16 16\ndt 0.1\n"
                  f"temperature 0.001\nloadparticles 1\ncoordinates {coords}\nvelocities {vels}\nseed 3\n")
    p = F.load_datamain(str(dm))
    assert p.np == 3  # the file's count replaces numberparticles
    assert p.coordinates == str(coords).encode() and p.velocities == str(vels).encode()
    assert F.bytes_needed(p) > 0
    o = O.Oracle(p)
    assert o.p.np == 3
    assert np.array_equal(o.particles(), np.array(q).T.ravel())
    assert np.array_equal(o.particle_velocities(), np.array(u).T.ravel())


def test_velocities_file_count_must_match(tmp_path):
    coords = tmp_path / "p.coor"
    vels = tmp_path / "p.vel"
    _write(coords, [(1.0, 0.0, 0.0), (2.0, 0.0, 0.0)])
    _write(vels, [(0.0, 0.0, 0.0)])
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=2, initial_positions=2, coordinates=str(coords),
                    velocities=str(vels))
    with pytest.raises(RuntimeError):
        O.Oracle(p)


def test_initfluid_options_in_the_oracle():
    """initfluid 1: sqrt(kT/(rho dV)) N(0,1) per face with the mean replaced by the background;
    2: vy = vy0 + 0.04 sin(2 pi z / lz); 3: vz = 0.004 sin(16 pi (z + dz/2) / lz); 4: vx = 0.425
    (src/initializeFluid.cpp:85-140), z the cell centre (src/createCells.cpp:66-69)."""
    cells, box = (16, 16, 32), (8.0, 8.0, 24.0)
    dV = 0.5 * 0.5 * 0.75
    kT, rho = 0.3, 1.7
    bg = (0.1, -0.2, 0.05)
    o = O.Oracle(make_params(SCHEME_INCOMPRESSIBLE, cells=cells, box=box, initfluid=1, temperature=kT,
                             densfluid=rho, background=bg, seed=5))
    v = o.velocity().reshape(3, -1)
    assert np.allclose(v.mean(axis=1), bg, atol=1e-14)
    var = v.var(axis=1)
    assert np.all(np.abs(var / (kT / (rho * dV)) - 1) < 0.05)
    assert abs(np.corrcoef(v[0], v[1])[0, 1]) < 0.03
    # the draws are the step-0 normals of stream 9 at each cell index
    n0 = O.normals4(5, 0, 9, 0, 0)
    raw0 = np.sqrt(kT / (rho * dV)) * n0[:3]
    raw = np.array([np.sqrt(kT / (rho * dV)) * O.normals4(5, 0, 9, 0, i)[:3] for i in range(v.shape[1])]).T
    assert np.allclose(v[:, 0], raw0 - raw.mean(axis=1) + np.array(bg), atol=1e-13)
    z = (np.arange(cells[2]) + 0.5) * 0.75 - 12.0
    zc = np.repeat(z, cells[0] * cells[1])
    o2 = O.Oracle(make_params(SCHEME_INCOMPRESSIBLE, cells=cells, box=box, initfluid=2, background=bg))
    v2 = o2.velocity().reshape(3, -1)
    assert np.allclose(v2[1], bg[1] + 0.04 * np.sin(2 * np.pi / 24.0 * zc), atol=1e-15)
    assert np.allclose(v2[0], bg[0]) and np.allclose(v2[2], bg[2])
    o3 = O.Oracle(make_params(SCHEME_INCOMPRESSIBLE, cells=cells, box=box, initfluid=3, background=bg))
    v3 = o3.velocity().reshape(3, -1)
    assert np.allclose(v3[2], 0.004 * np.sin(2 * np.pi / 24.0 * 8 * (zc + 0.375)), atol=1e-15)
    assert np.all(v3[:2] == 0.0)
    o4 = O.Oracle(make_params(SCHEME_INCOMPRESSIBLE, cells=cells, box=box, initfluid=4, background=bg))
    v4 = o4.velocity().reshape(3, -1)
    assert np.all(v4[0] == 0.425) and np.all(v4[1:] == 0.0)
