"""Binary state frames (maddening_amd/frames.py) against the reference's
format and tests (src/maddening/api/binary_encoder.py,
tests/api/test_binary_encoder.py): layout, sorted schema, offsets, round
trip, subscriptions. On the GPU the device-packed frame of an LBM node's
state must be byte-identical to the frame packed from host copies."""
import struct

import numpy as np
import pytest
import torch

from maddening_amd.frames import BinaryStateEncoder, decode_frame, field_view


@pytest.fixture
def sample_state():
    return {"ball": {"position": torch.tensor(5.0), "velocity": torch.tensor(-3.2)},
            "heat_rod": {"temperature": torch.ones(20) * 20.0},
            "table": {"position": torch.tensor(0.0)}}


def test_schema_layout_and_offsets(sample_state):
    enc = BinaryStateEncoder(sample_state)
    s = enc.schema()
    assert s["type"] == "schema" and s["compression"] == "none"
    assert s["total_floats"] == 23 and s["frame_bytes"] == 8 + 23 * 4
    assert [(f["node"], f["field"]) for f in s["fields"]] == [
        ("ball", "position"), ("ball", "velocity"), ("heat_rod", "temperature"),
        ("table", "position")]
    off = 0
    for f in s["fields"]:
        assert f["offset"] == off
        off += f["count"]


def test_round_trip_values(sample_state):
    enc = BinaryStateEncoder(sample_state)
    frame = enc.encode(1.234, sample_state)
    assert len(frame) == enc.frame_bytes
    assert struct.unpack_from("d", frame, 0)[0] == 1.234
    t, vals = decode_frame(frame, enc.schema())
    assert t == 1.234 and vals.dtype == np.float32
    assert field_view(vals, enc.schema(), "ball", "velocity") == np.float32(-3.2)
    assert np.all(field_view(vals, enc.schema(), "heat_rod", "temperature") == 20.0)


def test_subscriptions(sample_state):
    enc = BinaryStateEncoder(sample_state, fields={"ball": ["position", "nonexistent"],
                                                   "ghost": ["x"]})
    assert enc.total_floats == 1 and enc.subscription == {"ball": ["position", "nonexistent"],
                                                          "ghost": ["x"]}
    frame = enc.encode(2.5, sample_state)
    assert len(frame) == 12 and np.frombuffer(frame, np.float32, offset=8)[0] == 5.0
    assert BinaryStateEncoder(sample_state, fields={}).frame_bytes == 8
    assert BinaryStateEncoder({}).encode(0.5, {}) == struct.pack("d", 0.5)


def test_compression_options():
    with pytest.raises(ValueError, match="compression"):
        BinaryStateEncoder({}, compression="lzma")
    try:
        import zstandard  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError, match="zstandard"):
            BinaryStateEncoder({}, compression="zstd")


def test_lbm_velocity_subscription_size():
    """A velocity-only subscriber of a D3Q19 32^3 state gets 3 of 22 floats per cell."""
    n = 32 ** 3
    state = {"lbm": {"velocity": np.zeros((n, 3), np.float32),
                     "f": np.zeros((n, 19), np.float32)}}
    full = BinaryStateEncoder(state)
    sub = BinaryStateEncoder(state, fields={"lbm": ["velocity"]})
    assert sub.frame_bytes == 8 + n * 3 * 4
    assert 1 - sub.frame_bytes / full.frame_bytes > 0.80


@pytest.mark.gpu
def test_device_packed_frame_equals_host_frame():
    from maddening_amd.node import LBMNode
    node = LBMNode("lbm", 1e-3, grid_shape=(20, 16, 12), viscosity=0.1)
    st = node.initial_state()
    st = node.run(st, {"body_force": [1e-4, 0.0, 2e-5]}, 10)
    host = {"lbm": {k: v.cpu() for k, v in st.items()}}
    for fields in (None, {"lbm": ["velocity"]}, {"lbm": ["density", "f"]}):
        enc = BinaryStateEncoder({"lbm": st}, fields=fields)
        a = enc.encode(0.25, {"lbm": st})
        b = BinaryStateEncoder(host, fields=fields).encode(0.25, host)
        assert a == b
        a2 = enc.encode(0.5, {"lbm": st})          # staging buffers reused
        assert a2[8:] == a[8:]
    _, vals = decode_frame(a, enc.schema())
    got = field_view(vals, enc.schema(), "lbm", "f")
    assert np.array_equal(got, st["f"].cpu().numpy())       # reference (*G, Q) order
