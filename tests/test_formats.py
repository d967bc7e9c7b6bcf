"""On-disk formats (nefct/io/local_io_mixin.py, data_io_mixin.py,
json_io_mixin.py), on the CPU.  Follows the reference's own tests
(tests/nef_class/io/test_local_io_mixin.py, test_data_io_mixin.py): save and
load round trips of a flat and a nested object, dump_data / load_data, and
adds the JSON functions, deepdish's layout rules (dict / list / tuple / None
groups, scalars as attributes, unpacked non-dict values, partial loads),
and the file's structure as the HDF5 library's own dump tool reads it."""
import dataclasses
import os
import shutil
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")


@pytest.fixture()
def fmt(tmp_path, monkeypatch):
    from nefct_hip import formats
    try:
        formats.h5lib()
    except formats.FormatError as e:
        pytest.skip(str(e))
    monkeypatch.setenv("NEFCT_RESOURCE_DIR", str(tmp_path / "resources"))
    return formats


def _config():
    from nefct_hip.reference import DetectorConfig, ScannerConfig3D
    return ScannerConfig3D("flat", 1500.0, 1000.0, DetectorConfig(64, 96.0, 1.5), DetectorConfig(48, 72.0, -2.0))


def test_nested_dict_roundtrip_and_layout(fmt, tmp_path):
    data = {
        "vol": np.arange(24, dtype=np.float32).reshape(2, 3, 4),
        "counts": np.array([1, -2, 3], np.int64),
        "mask": np.array([True, False, True]),
        "n": 7, "x": 2.5, "flag": True, "name": "flat detector",
        "inner": {"a": np.ones((2, 2)), "b": None, "lst": [1, "two", np.zeros(3, np.int32)], "tup": (4.0, 5.0)},
    }
    path = str(tmp_path / "d.hdf5")
    fmt.h5_save(path, data)
    back = fmt.h5_load(path)
    assert set(back) == set(data)
    assert back["vol"].dtype == np.float32 and np.array_equal(back["vol"], data["vol"])
    assert back["counts"].dtype == np.int64 and np.array_equal(back["counts"], data["counts"])
    assert back["mask"].dtype == np.bool_ and np.array_equal(back["mask"], data["mask"])
    assert back["n"] == 7 and back["x"] == 2.5 and back["flag"] is True and back["name"] == "flat detector"
    inner = back["inner"]
    assert inner["b"] is None
    assert isinstance(inner["lst"], list) and inner["lst"][:2] == [1, "two"]
    assert np.array_equal(inner["lst"][2], np.zeros(3, np.int32))
    assert inner["tup"] == (4.0, 5.0)
    # partial loads of a group, a dataset and a scalar
    assert np.array_equal(fmt.h5_load(path, "/inner/a"), np.ones((2, 2)))
    assert fmt.h5_load(path, "/inner")["tup"] == (4.0, 5.0)
    assert fmt.h5_load(path, "/n") == 7


def test_non_dict_value_is_unpacked(fmt, tmp_path):
    path = str(tmp_path / "a.hdf5")
    a = np.random.default_rng(0).standard_normal((5, 6)).astype(np.float32)
    fmt.h5_save(path, a)
    back = fmt.h5_load(path)
    assert isinstance(back, np.ndarray) and np.array_equal(back, a)
    fmt.h5_save(path, [1.0, 2.0])
    assert fmt.h5_load(path) == [1.0, 2.0]


@pytest.mark.skipif(not (shutil.which("h5dump") or os.path.exists("/opt/conda/bin/h5dump")),
                    reason="h5dump not available")
def test_structure_as_the_hdf5_dump_tool_reads_it(fmt, tmp_path):
    """deepdish's layout, as h5dump (the HDF5 library's tool) sees the file:
    groups with PyTables titles, datasets, scalar attributes."""
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    path = str(tmp_path / "s.hdf5")
    fmt.h5_save(path, {"img": np.zeros((3, 4), np.float32), "k": 3, "sub": {"s": "txt"}, "none": None,
                       "l": [np.ones(2)]})
    out = subprocess.run([h5dump, "-H", path], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "img"' in out and "H5T_IEEE_F32LE" in out and "( 3, 4 )" in out
    assert 'GROUP "sub"' in out and 'ATTRIBUTE "s"' in out
    assert 'GROUP "none"' in out and 'GROUP "l"' in out and 'DATASET "i0"' in out
    assert 'ATTRIBUTE "k"' in out and 'ATTRIBUTE "TITLE"' in out
    titles = subprocess.run([h5dump, "-a", "/sub/TITLE", path], capture_output=True, text=True, check=True).stdout
    assert '"dict:1"' in titles
    titles = subprocess.run([h5dump, "-a", "/none/TITLE", path], capture_output=True, text=True, check=True).stdout
    assert '"nonetype"' in titles


def test_save_load_objects(fmt, tmp_path):
    """local_io_mixin: save/load of a flat and a nested object (the
    reference's test_local_io_mixin.py)."""
    from nefct_hip.reference import DetectorConfig, Image3D, ProjectionSequence3D, ScannerConfig3D
    det = DetectorConfig(64, 96.0, 1.5)
    p = det.save(str(tmp_path / "det.hdf5"))
    assert DetectorConfig.load(p) == det
    cfg = _config()
    p = cfg.save(str(tmp_path / "cfg.hdf5"))
    back = ScannerConfig3D.load(p)
    assert back == cfg and isinstance(back.detector_a, DetectorConfig)
    img = Image3D(torch.arange(60, dtype=torch.float32).reshape(5, 4, 3), (1.0, 2.0, 3.0), (10.0, 8.0, 6.0))
    p = img.save()                       # default path under the resource directory
    assert p.startswith(str(tmp_path / "resources")) and p.endswith(".hdf5")
    ib = Image3D.load(p)
    assert torch.equal(ib.data, img.data) and ib.center == img.center and ib.size == img.size
    assert ib.timestamps is None
    proj = ProjectionSequence3D(torch.ones((4, 3, 2)), cfg, np.array([0.0, 0.5]), np.array([1.0, -1.0]))
    pb = ProjectionSequence3D.load(proj.save(str(tmp_path / "proj.hdf5")))
    assert torch.equal(pb.data, proj.data) and pb.scanner == cfg
    assert np.array_equal(pb.angles, proj.angles) and np.array_equal(pb.offsets, proj.offsets)
    # a partial load reads one field's node
    assert ScannerConfig3D.load(p_cfg := str(tmp_path / "cfg.hdf5")) == cfg
    assert DetectorConfig.load(p_cfg, "/detector_b") == cfg.detector_b


def test_dump_and_load_data(fmt):
    """data_io_mixin: data goes to its own file, the field holds the path
    (the reference's test_data_io_mixin.py)."""
    from nefct_hip.reference import Image3D, ProjectionSequence3D
    img = Image3D(torch.rand((3, 4, 5)), (0.0, 0.0, 0.0), (3.0, 4.0, 5.0))
    dumped = img.dump_data()
    assert isinstance(dumped.data, str) and os.path.exists(dumped.data)
    back = dumped.load_data()
    assert torch.equal(back.data, img.data)
    proj = ProjectionSequence3D(torch.rand((2, 2, 2)), _config(), np.zeros(2))
    d2 = proj.dump_all_data()
    assert isinstance(d2.data, str)
    assert torch.equal(d2.load_all_data().data, proj.data)


def test_json(fmt, tmp_path):
    """json_io_mixin: objects whose data is not dumped are refused; the rest
    round-trips through a string and a file."""
    from nefct_hip.reference import Image3D, ScannerConfig3D
    cfg = _config()
    s = cfg.json_dumps()
    assert ScannerConfig3D.json_loads(s) == cfg
    assert ScannerConfig3D.json_load(cfg.json_dump(str(tmp_path / "c.json"))) == cfg
    img = Image3D(torch.zeros((2, 2, 2)), (0.0, 0.0, 0.0), (2.0, 2.0, 2.0))
    with pytest.raises(ValueError):
        img.json_dumps()
    d = img.dump_data()
    back = Image3D.json_loads(d.json_dumps())
    assert back.data == d.data
    assert torch.equal(back.load_data().data, img.data)


def test_dataclass_fields_cover_reference_keys():
    """The mirror classes carry the reference classes' field names
    (scanner_config.py:44-50, detector_config.py:9-12, image.py:46-51,
    projection.py:69-76), so dicts written by either side map field by field."""
    from nefct_hip.reference import DetectorConfig, Image3D, ProjectionSequence3D, ScannerConfig3D
    names = lambda c: [f.name for f in dataclasses.fields(c)]
    assert names(DetectorConfig) == ["number", "size", "offset"]
    assert names(ScannerConfig3D) == ["mode", "SID", "SAD", "detector_a", "detector_b"]
    assert names(Image3D)[:3] == ["data", "center", "size"] and "timestamps" in names(Image3D)
    assert names(ProjectionSequence3D)[:3] == ["data", "scanner", "angles"]
    assert {"offsets", "timestamps"} <= set(names(ProjectionSequence3D))
