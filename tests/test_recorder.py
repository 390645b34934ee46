"""Recording format (CPU): RingBuffer, the CBOR codec against RFC 8949's examples, and
ViewerFile round trips in `.json` and `.bin` with the reference's record layout
(newton/_src/viewer/viewer_file.py; scenarios of newton/tests/test_recorder.py:28-362)."""

from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from newton_sph_dev_amd import Axis, Model, ModelBuilder, cbor
from newton_sph_dev_amd.recorder import RingBuffer, ViewerFile


def test_ringbuffer_keeps_the_last_items():
    rb = RingBuffer(3)
    assert len(rb) == 0 and rb.to_list() == []
    for x in "abc":
        rb.append(x)
    assert rb.to_list() == ["a", "b", "c"] and rb[0] == "a" and rb[2] == "c"
    rb.append("d")
    assert len(rb) == 3 and rb.to_list() == ["b", "c", "d"] and rb[0] == "b"
    rb.append("e")
    assert rb.to_list() == ["c", "d", "e"]
    rb[1] = "D"
    assert list(rb) == ["c", "D", "e"]


def test_ringbuffer_edge_cases():
    rb = RingBuffer(2)
    with pytest.raises(IndexError):
        _ = rb[0]
    with pytest.raises(IndexError):
        rb[0] = "x"
    assert list(rb) == []
    rb.append("x")
    rb.append("y")
    rb.append("z")
    assert list(rb) == ["y", "z"]
    rb.clear()
    assert len(rb) == 0
    rb.from_list(["1", "2", "3", "4"])
    assert rb.to_list() == ["3", "4"]


# RFC 8949 Appendix A (examples of encoded CBOR data items)
RFC_VECTORS = [
    (0, "00"), (1, "01"), (10, "0a"), (23, "17"), (24, "1818"), (25, "1819"), (100, "1864"),
    (1000, "1903e8"), (1000000, "1a000f4240"), (1000000000000, "1b000000e8d4a51000"),
    (18446744073709551615, "1bffffffffffffffff"), (-1, "20"), (-10, "29"), (-100, "3863"),
    (-1000, "3903e7"), (1.1, "fb3ff199999999999a"), (1.0e300, "fb7e37e43c8800759c"),
    (-4.1, "fbc010666666666666"), (False, "f4"), (True, "f5"), (None, "f6"),
    (b"", "40"), (b"\x01\x02\x03\x04", "4401020304"), ("", "60"), ("a", "6161"), ("IETF", "6449455446"),
    ("ü", "62c3bc"), ("水", "63e6b0b4"), ([], "80"), ([1, 2, 3], "83010203"),
    ([1, [2, 3], [4, 5]], "8301820203820405"), ({}, "a0"), ({"a": 1, "b": [2, 3]}, "a26161016162820203"),
    (["a", {"b": "c"}], "826161a161626163"),
]


@pytest.mark.parametrize("value,hexstr", RFC_VECTORS)
def test_cbor_matches_rfc8949_examples(value, hexstr):
    assert cbor.dumps(value).hex() == hexstr
    assert cbor.loads(bytes.fromhex(hexstr)) == value


def test_cbor_decodes_other_rfc_forms():
    # half and single precision floats, indefinite lengths and a tag (RFC 8949 Appendix A)
    assert cbor.loads(bytes.fromhex("f93c00")) == 1.0
    assert cbor.loads(bytes.fromhex("f97bff")) == 65504.0
    assert cbor.loads(bytes.fromhex("fa47c35000")) == 100000.0
    assert cbor.loads(bytes.fromhex("9f018202039f0405ffff")) == [1, [2, 3], [4, 5]]
    assert cbor.loads(bytes.fromhex("bf61610161629f0203ffff")) == {"a": 1, "b": [2, 3]}
    assert cbor.loads(bytes.fromhex("5f42010243030405ff")) == b"\x01\x02\x03\x04\x05"
    assert cbor.loads(bytes.fromhex("c11a514b67b0")) == 1363896240
    with pytest.raises(ValueError):
        cbor.loads(bytes.fromhex("8301"))


def _model():
    b = ModelBuilder(up_axis="Y")
    b.add_particle_grid(pos=(0.0, 0.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.1, 0.0, 0.0), dim_x=3, dim_y=2,
                        dim_z=2, cell_x=0.1, cell_y=0.1, cell_z=0.1, mass=0.5, jitter=0.01, radius_mean=0.05)
    return b.finalize("cpu")


def _states(m, n=3):
    out = []
    for i in range(n):
        s = m.state()
        s.particle_q += float(i)
        s.particle_qd.fill_(0.25 * i)
        s.particle_f.fill_(-float(i))
        out.append(s)
    return out


@pytest.mark.parametrize("ext", [".json", ".bin"])
def test_model_and_state_round_trip(tmp_path, ext):
    m = _model()
    states = _states(m)
    rec = ViewerFile(str(tmp_path / "unused.json"), auto_save=False, device="cpu")
    rec.record_model(m)
    for s in states:
        rec.record(s)
    path = str(tmp_path / f"rec{ext}")
    rec.save_recording(path)
    back = ViewerFile(path, auto_save=False, device="cpu")
    back.load_recording()
    assert back.has_model() and back.get_frame_count() == 3
    restored = Model("cpu")
    back.load_model(restored)
    assert restored.particle_count == m.particle_count
    assert torch.equal(restored.particle_q, m.particle_q)
    assert torch.equal(restored.particle_mass, m.particle_mass)
    assert torch.equal(restored.particle_flags, m.particle_flags)
    assert restored.up_axis == Axis.Y and restored.particle_max_radius == m.particle_max_radius
    for k, s in enumerate(states):
        t = restored.state()
        back.load_state(t, k)
        for name in ("particle_q", "particle_qd", "particle_f"):
            assert torch.equal(getattr(t, name), getattr(s, name)), (k, name)


def test_json_layout_is_the_viewerfile_layout(tmp_path):
    m = _model()
    rec = ViewerFile(str(tmp_path / "r.json"), auto_save=False, device="cpu")
    rec.set_model(m)
    rec.log_state(m.state())
    rec.save_recording()
    tree = json.load(open(tmp_path / "r.json"))
    assert tree["__type__"] == "dict" and set(tree["items"]) == {"model", "states"}
    model = tree["items"]["model"]
    assert model["__type__"] == "Model" and "attributes" in model
    q = model["attributes"]["particle_q"]
    assert q["__type__"] == "warp.array" and q["__dtype__"] == "<class 'warp.types.vec3f'>"
    assert q["data"]["dtype"] == "float32" and q["data"]["shape"] == [m.particle_count, 3]
    assert np.array_equal(np.array(json.loads(q["data"]["data"]), np.float32), m.particle_q.numpy())
    assert model["attributes"]["particle_mass"]["__dtype__"] == "<class 'warp.types.float32'>"
    assert model["attributes"]["particle_flags"]["__dtype__"] == "<class 'warp.types.int32'>"
    assert model["attributes"]["particle_grid"] == {"__type__": "warp.HashGrid", "data": None}
    frame = tree["items"]["states"]["items"][0]
    assert frame["__type__"] == "dict" and frame["items"]["particle_q"]["__type__"] == "warp.array"


def test_repeated_array_is_written_once(tmp_path):
    m = _model()
    s = m.state()
    rec = ViewerFile(str(tmp_path / "r.bin"), auto_save=False, device="cpu")
    rec.history.append({"particle_q": s.particle_q})
    rec.history.append({"particle_q": s.particle_q})
    rec.save_recording()
    tree = cbor.loads(open(tmp_path / "r.bin", "rb").read())
    frames = tree["items"]["states"]["items"]
    assert frames[0]["items"]["particle_q"]["__type__"] == "warp.array"
    assert frames[1]["items"]["particle_q"]["__type__"] == "warp.array_ref"
    back = ViewerFile(str(tmp_path / "r.bin"), auto_save=False, device="cpu")
    back.load_recording()
    assert torch.equal(back.history[1]["particle_q"], s.particle_q)


def test_ring_buffer_history_saves_the_last_frames(tmp_path):
    m = _model()
    rec = ViewerFile(str(tmp_path / "r.json"), auto_save=False, max_history_size=3, device="cpu")
    rec.record_model(m)
    states = _states(m, 5)
    for s in states:
        rec.record(s)
    assert len(rec.history) == 3
    rec.save_recording()
    back = ViewerFile(str(tmp_path / "r.json"), auto_save=False, max_history_size=5, device="cpu")
    back.load_recording()
    assert len(back.history) == 3
    assert torch.equal(back.history[0]["particle_q"], states[2].particle_q)


def test_unsupported_extension(tmp_path):
    rec = ViewerFile(str(tmp_path / "r.txt"), auto_save=False, device="cpu")
    with pytest.raises(ValueError):
        rec.save_recording()
