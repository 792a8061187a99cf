"""Serialization savepoints (the analogue of the reference's serialbox calls,
advection_CEtimestep_twostep_GPUBC.f90:122-142, 355-375): host-memory fields
through the C ABI (no GPU needed), the .npy files read back with numpy's
pickle-free loader, the index, and the argument checks."""
import numpy as np
import pytest


@pytest.fixture()
def ser(tmp_path):
    from advectionhpctester_amd import serialization
    d = tmp_path / "ser"
    with serialization.Serializer(d) as s:
        yield s, d, serialization


def test_host_fields_round_trip(ser):
    s, d, S = ser
    rng = np.random.default_rng(1)
    a = rng.standard_normal((5, 4, 3))          # C-ordered (n3, n2, n1): Fortran (3, 4, 5)
    b = rng.standard_normal((7, 2)).astype(np.float32)
    s.write("mysnapshot.run-in", "th", a)
    s.write("mysnapshot.run-in", "th_lr", b)
    s.write("mysnapshot.run-out", "th", a * 2)
    got = S.read(d)
    assert sorted(got) == ["mysnapshot.run-in", "mysnapshot.run-out"]
    assert got["mysnapshot.run-in"]["th"].shape == (3, 4, 5)
    # the Fortran array (n1, n2, n3) is the transpose of the C-ordered one, same bytes
    assert np.array_equal(got["mysnapshot.run-in"]["th"], a.T)
    assert got["mysnapshot.run-in"]["th_lr"].dtype == np.float32
    assert np.array_equal(got["mysnapshot.run-in"]["th_lr"][:, :, 0], b.T)
    assert np.array_equal(got["mysnapshot.run-out"]["th"], 2 * a.T)
    assert np.asfortranarray(got["mysnapshot.run-in"]["th"]).tobytes(order="F") == a.tobytes()
    lines = (d / "savepoints.txt").read_text().splitlines()
    assert lines[0].split() == ["mysnapshot.run-in", "th", "mysnapshot.run-in__th.npy", "f8",
                                "3", "4", "5"]


def test_rewrite_keeps_last_version(ser):
    s, d, S = ser
    s.write("sp", "x", np.zeros((2, 2, 2)))
    s.write("sp", "x", np.ones((2, 2, 2)))
    assert np.array_equal(S.read(d)["sp"]["x"], np.ones((2, 2, 2)))


def test_bad_arguments(ser):
    s, d, S = ser
    from advectionhpctester_amd._lib import MpdataError
    with pytest.raises(MpdataError):
        s.write("bad name", "x", np.zeros((2, 2, 2)))
    with pytest.raises(MpdataError):
        s.write("sp", "../x", np.zeros((2, 2, 2)))
    with pytest.raises(TypeError):
        s.write("sp", "x", np.zeros((2, 2, 2), dtype=np.int32))


def test_no_serializer_open(tmp_path):
    import ctypes
    from advectionhpctester_amd._lib import LIB
    LIB.mpdata_serializer_close()
    a = np.zeros(8)
    rc = LIB.mpdata_serialize_field_f64(b"sp", b"x", ctypes.c_void_p(a.ctypes.data), 2, 2, 2, None)
    assert rc != 0 and b"no serializer" in LIB.mpdata_last_error()
