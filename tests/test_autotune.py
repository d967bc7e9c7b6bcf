"""Autotuner cache (CPU) and tuned adjoint (GPU): the tuned variant gives the
same result as the default one (identical weights, different work mapping)."""
import json
import math

import numpy as np
import pytest


def test_views_bucket_and_cache_roundtrip(tmp_path, monkeypatch):
    from nefct_hip import autotune as at
    assert [at._views_bucket(n) for n in (1, 2, 3, 64, 65, 720)] == [1, 2, 4, 64, 128, 1024]
    monkeypatch.setenv("NEFCT_AUTOTUNE_CACHE", str(tmp_path / "c" / "tune.json"))
    monkeypatch.setattr(at, "_cache", None)
    at._store("k1", {"flags": 1024, "variant": "x"})
    monkeypatch.setattr(at, "_cache", None)
    assert at._load()["k1"]["flags"] == 1024
    assert json.load(open(tmp_path / "c" / "tune.json"))["k1"]["variant"] == "x"


def test_candidates_are_valid_flag_sets():
    from nefct_hip import _lib
    from nefct_hip.autotune import ADJOINT_CANDIDATES
    allowed = (_lib.F_TILE_MASK | _lib.F_LANE_STRIDE2 | _lib.F_ADJ_FLOAT_ATOMIC | _lib.F_ADJ_WINDOW)
    for name, fl in ADJOINT_CANDIDATES.items():
        assert fl & ~allowed == 0, name
    assert len(set(ADJOINT_CANDIDATES.values())) == len(ADJOINT_CANDIDATES)


def test_forward_candidates_are_valid_flag_sets():
    from nefct_hip import _lib
    from nefct_hip.autotune import FORWARD_CANDIDATES
    allowed = _lib.F_FWD_TILE_MASK | _lib.F_FWD_VG_MASK
    for name, fl in FORWARD_CANDIDATES.items():
        assert fl & ~allowed == 0, name
        assert 1 <= (fl & _lib.F_FWD_VG_MASK) >> _lib.F_FWD_VG_SHIFT <= 6, name
    assert len(set(FORWARD_CANDIDATES.values())) == len(FORWARD_CANDIDATES) == 16
    with pytest.raises(ValueError):
        _lib.fwd_vg(7)


@pytest.mark.gpu
def test_tuned_forward_matches_default(tmp_path, monkeypatch):
    """Every forward tile and view group gives the default's result
    bit for bit (each ray is walked by one lane with the same arithmetic),
    and the tuned choice is cached."""
    torch = pytest.importorskip("torch")
    import nefct_hip as nh
    from nefct_hip import autotune as at
    from nefct_hip.phantom import ellipsoid_phantom
    monkeypatch.setenv("NEFCT_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(at, "_cache", None)
    vol = nh.VolumeGeometry.cube((40, 36, 28), 1.5)
    det = nh.ConeGeometry.flat(200.0, 300.0, 70, 1.4, 45, 1.4)
    ang = torch.tensor((2 * math.pi * np.arange(40) / 40).astype(np.float32))
    x = ellipsoid_phantom(vol, 5, 6).cuda()
    ref = nh.project(x, vol, det, ang)
    for name, fl in at.FORWARD_CANDIDATES.items():
        assert torch.equal(nh.project(x, vol, det, ang, flags=fl), ref), name
        got = nh.project(x, vol, det, ang, flags=nh.REF_FORWARD | fl)
        assert torch.equal(got, nh.project(x, vol, det, ang, flags=nh.REF_FORWARD)), name
    flags, rec = at.tune_forward(vol, det, ang, sample_views=32, reps=1)
    assert rec["variant"] in at.FORWARD_CANDIDATES
    assert torch.equal(at.project_tuned(x, vol, det, ang), ref)
    flags2, rec2 = at.tune_forward(vol, det, ang)
    assert flags2 == flags and rec2 == rec
    assert at.tuned_flags(vol, det, 40, 0, "forward") == flags


@pytest.mark.gpu
def test_tuned_adjoint_matches_default(tmp_path, monkeypatch):
    torch = pytest.importorskip("torch")
    import nefct_hip as nh
    from nefct_hip import autotune as at
    monkeypatch.setenv("NEFCT_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(at, "_cache", None)
    vol = nh.VolumeGeometry.cube((40, 36, 28), 1.5)
    det = nh.ConeGeometry.flat(200.0, 300.0, 64, 1.4, 48, 1.4)
    ang = torch.tensor((2 * math.pi * np.arange(30) / 30).astype(np.float32))
    y = torch.rand((30, 48, 64), device="cuda")
    flags, rec = at.tune_adjoint(vol, det, ang, sample_views=16, reps=1)
    assert rec["variant"] in at.ADJOINT_CANDIDATES
    assert set(rec["ms_on_sample"]) == set(at.ADJOINT_CANDIDATES)
    ref = nh.backproject(y, vol, det, ang)
    got = at.backproject_tuned(y, vol, det, ang)
    assert torch.allclose(got, ref, rtol=1e-5, atol=1e-5 * ref.abs().max().item())
    # cached: a second call does not re-time
    flags2, rec2 = at.tune_adjoint(vol, det, ang)
    assert flags2 == flags and rec2 == rec
    # reference-mode flags are kept in the tuned set
    fr, _ = at.tune_adjoint(vol, det, ang, mode_flags=nh.REF_ADJOINT, sample_views=8, reps=1)
    assert fr & nh.REF_ADJOINT == nh.REF_ADJOINT
