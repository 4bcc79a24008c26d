"""HDF5 output in the reference's layout (lib/io/fileio.cc): dataset names,
FP64 round trip, append, errors.  Runs on the CPU (no simulation needed)."""
import numpy as np
import pytest

from luxum_amd._lib import LuxumError
from luxum_amd.sim import h5_dataset_name, h5_read, h5_write


def test_dataset_names_follow_the_reference_ostream():
    # fileio.cc:175 builds the name with  temp << "t=" << t  (t a double,
    # default ostream formatting: %g with 6 significant digits)
    for t, name in [(0, "t=0"), (3, "t=3"), (100, "t=100"), (999999, "t=999999"), (1000000, "t=1e+06"),
                    (1234567, "t=1.23457e+06"), (0.5, "t=0.5"), (2000, "t=2000")]:
        assert h5_dataset_name(t) == name


def test_round_trip_append_and_errors(tmp_path):
    path = str(tmp_path / "Ez_data.h5")
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, 4, 3))
    b = rng.standard_normal((5, 4, 3))
    h5_write(path, 0, a)
    h5_write(path, 3, b, append=True)
    assert np.array_equal(h5_read(path, "t=0"), a)
    assert np.array_equal(h5_read(path, "t=3"), b)
    with pytest.raises(LuxumError):
        h5_write(path, 3, b, append=True)          # dataset exists
    with pytest.raises(LuxumError):
        h5_read(path, "t=6")                       # no such dataset
    with pytest.raises(LuxumError):
        h5_read(str(tmp_path / "missing.h5"), "t=0")
    h5_write(path, 6, a)                           # truncates
    with pytest.raises(LuxumError):
        h5_read(path, "t=0")
    assert np.array_equal(h5_read(path, "t=6"), a)


def test_named_dataset(tmp_path):
    # material inputs of the 3D CPML solver carry one dataset named "data"
    # (maxwell_three_dim_pml.cc:116-118)
    path = str(tmp_path / "airplane_100_x.h5")
    a = np.arange(24, dtype=np.float64).reshape(2, 3, 4) * 1e-6
    h5_write(path, "data", a)
    assert np.array_equal(h5_read(path, "data"), a)
    with pytest.raises(LuxumError):
        h5_read(path, "t=0")


def test_large_datasets_round_trip(tmp_path):
    # snapshot-sized datasets appended to one file read back dataset by dataset
    path = str(tmp_path / "Ez_data.h5")
    rng = np.random.default_rng(5)
    arrs = [rng.standard_normal((64, 80, 101)) for _ in range(3)]   # 4.1 MB each
    arrs.append(rng.standard_normal((130, 128, 127)))               # 16.9 MB
    for q, a in enumerate(arrs):
        h5_write(path, 3 * q, a, append=q > 0)
    for q, a in enumerate(arrs):
        assert np.array_equal(h5_read(path, f"t={3 * q}"), a)
