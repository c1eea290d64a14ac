"""State interchange in the reference's checkpoint format
(core/simulation/checkpoint.py:69-199, 272-330)."""
import json

import numpy as np
import pytest

from maddening_amd import checkpoint as ck


def test_save_writes_reference_keys_and_manifest(tmp_path):
    rng = np.random.default_rng(0)
    st = {"f": rng.random((4, 3, 9)).astype(np.float32),
          "density": np.ones((4, 3), np.float32),
          "wall_mask": np.zeros((4, 3), np.uint8)}
    out = ck.save_state({"lbm": st}, tmp_path / "snap", extra={"step": 7})
    assert out.suffix == ".npz"
    with np.load(out, allow_pickle=False) as d:
        assert sorted(d.files) == ["lbm/density", "lbm/f", "lbm/wall_mask"]
        np.testing.assert_array_equal(d["lbm/f"], st["f"])
    man = json.loads(out.with_suffix(".npz.manifest.json").read_text())
    assert man["schema_version"] == 1 and man["extra"] == {"step": 7}
    assert man["sha256"] == ck.file_sha256(out) and man["size_bytes"] == out.stat().st_size
    ck.verify_manifest(out)
    nodes, meta = ck.load_arrays(out)
    assert set(nodes) == {"lbm"} and meta == {}
    np.testing.assert_array_equal(nodes["lbm"]["density"], st["density"])


def test_manifest_detects_tampering(tmp_path):
    out = ck.save_state({"a": {"x": np.arange(5)}}, tmp_path / "s.npz")
    with open(out, "ab") as fh:
        fh.write(b"\0")
    with pytest.raises(ck.CheckpointIntegrityError):
        ck.verify_manifest(out)


def test_bad_key_format_rejected(tmp_path):
    p = tmp_path / "bad.npz"
    np.savez(p, nokey=np.zeros(2))
    with pytest.raises(ValueError, match="Unexpected key format"):
        ck.load_arrays(p)


@pytest.mark.gpu
def test_node_round_trip_continues_identically(tmp_path):
    import torch
    from maddening_amd.node import LBMNode
    rng = np.random.default_rng(1)
    wall = rng.random((10, 8, 12)) < 0.1
    node = LBMNode(name="lbm", timestep=1e-3, grid_shape=(10, 8, 12), wall_mask=wall)
    bi = {"body_force": [0.0, 0.0, 1e-5]}
    st = node.run(node.initial_state(), bi, 5)
    out = ck.save_state({"lbm": st}, tmp_path / "snap")
    with np.load(out, allow_pickle=False) as d:
        assert d["lbm/f"].shape == (10, 8, 12, 19) and d["lbm/f"].dtype == np.float32
    back = ck.load_state(out, {"lbm": node}, verify=True)["lbm"]
    for key in st:
        assert torch.equal(back[key].cpu(), st[key].cpu()), key
    a = node.run(st, bi, 4)
    b = node.run(back, bi, 4)
    assert torch.equal(a["f"].cpu(), b["f"].cpu())
    with pytest.raises(ValueError, match="node mismatch"):
        ck.load_state(out, {"other": node})

