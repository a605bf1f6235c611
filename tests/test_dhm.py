"""DHM field files (include/dohp_dhm.h): the layout of dohp's viewer (src/viewer/dhm.c, contviewdhm.c) written
with HDF5 and read back.  Host-only: the library needs no GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from dohp_amd.mesh import block_vertices, sine_modes_state

D = pytest.importorskip("dohp_amd.dhm")


def _h5dump():
    for p in (shutil.which("h5dump"), "/opt/conda/bin/h5dump"):
        if p and os.path.exists(p):
            return p
    return None


def test_round_trip_spaces_steps_vectors(tmp_path):
    path = str(tmp_path / "fields.dhm")
    M, p = (3, 2, 4), 3
    V = block_vertices(M, box=(0, 1, 0, 1, 0, 1), perturb=0.1, seed=5)
    n = (3 * 3 + 1) * (3 * 2 + 1) * (3 * 4 + 1)
    u0 = sine_modes_state(np.random.default_rng(0).uniform(size=(n, 3)), 3)
    with D.DHM(path, D.WRITE) as f:
        f.set_time_units("s", 1.0)
        f.write_block_space("u-space", M, p, V, dirichlet_mask=5, fields=["u"])
        f.set_time(0.0)
        f.write_vec("u-space", "u", u0)
        f.write_vec("u-space", "residual", 2 * u0)
        f.set_time(0.5)
        f.write_vec("u-space", "u", u0 + 1)
        f.set_time(0.5)  # the same time stays in the same step
        f.write_vec("u-space", "w", -u0, unit_scale=3.0)
    with D.DHM(path, D.READ) as f:
        np.testing.assert_array_equal(f.steps(), [0.0, 0.5])
        sp = f.read_block_space("u-space")
        assert sp["M"] == M and sp["degree"] == p and sp["dirichlet_mask"] == 5 and sp["bs"] == 1
        np.testing.assert_array_equal(sp["vcoords"], V)
        f.set_step(0)
        np.testing.assert_array_equal(f.read_vec("u", n), u0)
        np.testing.assert_array_equal(f.read_vec("residual", n), 2 * u0)
        f.set_step(1)
        np.testing.assert_array_equal(f.read_vec("u", n), u0 + 1)
        np.testing.assert_array_equal(f.read_vec("w", n), -u0)
        with pytest.raises(D.DHMError):
            f.read_vec("u", n + 1)
        with pytest.raises(D.DHMError):
            f.set_step(2)
        with pytest.raises(D.DHMError):
            f.read_block_space("nope")
    # append: a later step keeps the earlier ones
    with D.DHM(path, D.APPEND) as f:
        f.set_time(1.0)
        f.write_vec("u-space", "u", u0 + 2)
    with D.DHM(path, D.READ) as f:
        np.testing.assert_array_equal(f.steps(), [0.0, 0.5, 1.0])
        f.set_step(2)
        np.testing.assert_array_equal(f.read_vec("u", n), u0 + 2)


def test_blocked_vector_and_layout(tmp_path):
    path = str(tmp_path / "blocked.dhm")
    M = (1, 1, 1)
    V = block_vertices(M)
    x = np.arange(24, dtype=np.float64)
    with D.DHM(path, D.WRITE) as f:
        f.write_block_space("xyz", M, 1, V, fields=["x", "y", "z"])
        f.set_time(2.0)
        f.write_vec("xyz", "coords", x, bs=3)
        with pytest.raises(D.DHMError):
            f.write_vec("xyz", "bad", x[:23], bs=3)
        with pytest.raises(D.DHMError):
            f.write_vec("missing-space", "v", x)
    with D.DHM(path, D.READ) as f:
        assert f.read_block_space("xyz", with_coords=False)["bs"] == 3
        np.testing.assert_array_equal(f.read_vec("coords", 24, bs=3), x)
    h5 = _h5dump()
    if h5 is None:
        pytest.skip("h5dump not available")
    out = subprocess.run([h5, "-H", path], capture_output=True, text=True).stdout
    # the reference's groups, committed types and records (dhm.c:251-302, 142-210; contviewdhm.c:132-206)
    for token in ('GROUP "dohp"', 'GROUP "mesh"', 'GROUP "fs"', 'GROUP "step"', 'GROUP "type"', 'DATATYPE "FS_type"',
                  'DATATYPE "Vec_type"', 'DATATYPE "Time_type"', 'DATATYPE "Units_type"', 'DATATYPE "Field_type"',
                  'DATATYPE "BoundingBox_type"', 'GROUP "000"', 'DATASET "coords"', 'ATTRIBUTE "meta"',
                  'ATTRIBUTE "time"', 'DATASET "xyz"', '"number_of_subdomains"', '"ordered_subdomain"',
                  'DATASPACE  SIMPLE { ( 8, 3 ) / ( 8, 3 ) }'):
        assert token in out, token
