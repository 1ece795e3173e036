"""G4 step input: the HDF5 reader (lighthit_amd.hdf5) and G4Shower.from_hdf5.

The reader is pinned on files that libhdf5 itself wrote, whose contents their
own test suite documents (PyTables' compatibility files, present in this image
under /opt/conda): numeric arrays of both byte orders, a padded compound, a
contiguous compound with an array and a string member inside a nested group,
and a chunked compound with nested array members. The G4 fixture
(tests/golden/g4_steps.h5, written by make_g4_fixture.py) then checks the step
conversion of ``sources.py:362-393`` against a row-by-row restatement."""
from pathlib import Path

import numpy as np
import pytest

from lighthit_amd.hdf5 import File, list_datasets, read_dataset
from lighthit_amd.sources import FINE_STRUCTURE, G4Shower

GOLDEN = Path(__file__).resolve().parent / "golden"
LIBHDF5 = Path("/opt/conda/lib/python3.9/site-packages/tables/tests")
written_by_libhdf5 = pytest.mark.skipif(not LIBHDF5.is_dir(),
                                        reason="libhdf5-written files not in this image")


@written_by_libhdf5
@pytest.mark.parametrize("name, dtype", [("smpl_i32be.h5", ">i4"), ("smpl_i32le.h5", "<i4"),
                                         ("smpl_i64be.h5", ">i8"), ("smpl_i64le.h5", "<i8"),
                                         ("smpl_f64be.h5", ">f8"), ("smpl_f64le.h5", "<f8")])
def test_numeric_arrays_written_by_libhdf5(name, dtype):
    data = read_dataset(LIBHDF5 / name, "TestArray")
    assert data.dtype == np.dtype(dtype) and data.shape == (6, 5)
    np.testing.assert_array_equal(data, np.add.outer(np.arange(6), np.arange(5)))


@written_by_libhdf5
def test_padded_compound_written_by_libhdf5():
    data = read_dataset(LIBHDF5 / "itemsize.h5", "Test")
    assert data.dtype.itemsize == 16 and data.dtype.fields["B"][1] == 4
    assert data["A"].tolist() == [1, 2, 3] and data["B"].tolist() == [11, 12, 13]


@written_by_libhdf5
def test_contiguous_compound_in_a_nested_group_written_by_libhdf5():
    path = LIBHDF5 / "non-chunked-table.h5"
    assert list_datasets(path) == ["test_var"]
    data = File(path)["test_var/structure variable"]
    assert data.dtype.names == ("a", "b", "c", "d")
    assert np.all(data["a"] == 3.0) and np.all(data["b"] == 4.0)
    assert np.all(data["c"] == [2.0, 3.0]) and np.all(data["d"] == b"d")


@written_by_libhdf5
def test_chunked_compound_with_array_members_written_by_libhdf5():
    data = read_dataset(LIBHDF5 / "smpl_compound_chunked.h5", "CompoundChunked")
    m = np.arange(len(data))
    assert len(data) == 6
    np.testing.assert_array_equal(data["a_name"], m)
    assert np.all(data["c_name"] == b"Hello!")
    np.testing.assert_array_equal(data["d_name"], m[:, None, None] + np.add.outer(
        np.arange(5), np.arange(10))[None])
    np.testing.assert_allclose(data["e_name"], m * 0.96, rtol=0, atol=1e-6)
    np.testing.assert_allclose(data["f_name"], np.repeat((m * 1024.9637)[:, None], 10, 1))
    assert np.all(data["g_name"] == ord("m"))


@written_by_libhdf5
def test_compressed_datasets_are_refused():
    path = LIBHDF5 / "Tables_lzo2.h5"
    names = []

    def walk(group):
        for key in File(path).keys(group):
            child = f"{group}/{key}".lstrip("/")
            try:
                File(path).keys(child)
                walk(child)
            except ValueError:
                names.append(child)
    walk("")
    refused = 0
    for name in names:
        try:
            File(path)[name]
        except ValueError as error:
            refused += "filter" in str(error)
    assert refused > 0


def test_not_hdf5_is_refused(tmp_path):
    path = tmp_path / "plain.bin"
    path.write_bytes(b"not an hdf5 file at all")
    with pytest.raises(ValueError, match="not an HDF5"):
        File(path)


def test_g4_fixture_rows_read_back():
    rows = np.load(GOLDEN / "g4_steps_rows.npz")
    assert list_datasets(GOLDEN / "g4_steps.h5") == ["event_0", "event_1"]
    for key in ("event_0", "event_1"):
        got = read_dataset(GOLDEN / "g4_steps.h5", f"{key}/tracks")
        assert got.dtype == rows[key].dtype
        np.testing.assert_array_equal(got, rows[key])


def _steps_row_by_row(rows, reference_phase_index=1.35):
    """The step rule of sources.py:362-393 restated one row pair at a time."""
    out = []
    for j in range(1, len(rows)):
        a, b = rows[j - 1], rows[j]
        if a["uid"] != b["uid"]:
            continue
        start = np.array([a["x_m"], a["y_m"], a["z_m"]])
        chord = np.array([b["x_m"], b["y_m"], b["z_m"]]) - start
        length = float(np.sqrt(chord @ chord))
        beta = (a["beta"] + b["beta"]) / 2
        path = b["step_length_m"]
        if length > 0 and path > 0 and beta * reference_phase_index > 1:
            c0 = 2 * np.pi * FINE_STRUCTURE * 1e9 * path
            out.append((start, chord / length, length, c0, -c0 / beta ** 2, beta, a["t_ns"],
                        b["t_ns"], j, b["uid"]))
    return out


@pytest.mark.parametrize("event", [0, 1])
def test_g4_steps_follow_the_step_rule(event):
    rows = np.load(GOLDEN / "g4_steps_rows.npz")[f"event_{event}"]
    shower = G4Shower.from_hdf5(GOLDEN / "g4_steps.h5", event)
    e = shower.elements
    want = _steps_row_by_row(rows)
    assert 0 < len(want) == len(e) < len(rows) - 1          # cuts did drop steps
    for k, (start, direction, length, c0, c2, beta, t0, t1, row, uid) in enumerate(want):
        np.testing.assert_allclose(e.start_m[k], start, rtol=0, atol=1e-15)
        np.testing.assert_allclose(e.direction[k], direction, rtol=1e-14)
        assert e.length_m[k] == pytest.approx(length, rel=1e-14)
        assert e.coefficient0[k] == pytest.approx(c0, rel=1e-14)
        assert e.coefficient2[k] == pytest.approx(c2, rel=1e-14)
        assert e.beta[k] == pytest.approx(beta, rel=1e-15)
        assert (e.start_ns[k], e.end_ns[k], e.row_index[k], e.uid[k]) == (t0, t1, row, uid)
    assert shower.event == event and shower.input_path.endswith("g4_steps.h5")
    assert e.provenance["rows"] == len(rows)
    first = G4Shower.from_hdf5(GOLDEN / "g4_steps.h5", event, max_elements=4)
    assert len(first.elements) <= 4
    np.testing.assert_array_equal(first.elements.row_index, e.row_index[:len(first.elements)])


def test_g4_tracks_need_the_step_columns(tmp_path):
    with pytest.raises(KeyError):
        G4Shower.from_hdf5(GOLDEN / "g4_steps.h5", 7)
