"""The vessel voxelizer (lbm_geometry.py:17-142), CPU: the reference's own
tests (tests/nodes/test_lbm.py:460-492), closed-form distances, and the
product against the scalar oracle bit for bit."""
import numpy as np
import pytest

from maddening_amd.geometry import voxelize_vessel
from oracle import lbm_geometry_loops as og

REF = {"parent_radius": 6.0, "daughter_radius": 4.0, "parent_length": 30.0,
       "daughter_length": 25.0, "bifurcation_angle": 30.0}


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_reference_fluid_fraction_and_parent_centre(impl):
    """test_lbm.py:461-492: (64, 32, 32), fluid fraction in (5 %, 50 %),
    the parent tube's centre (15, 15, 15) is fluid."""
    mask = voxelize_vessel((64, 32, 32), REF) if impl == "product" else \
        og.vessel_mask((64, 32, 32), REF)
    assert mask.shape == (64, 32, 32) and mask.dtype == np.bool_
    frac = float((~mask).sum()) / mask.size
    assert 0.05 < frac < 0.50, frac
    assert not mask[15, 15, 15]


def _closed_form_checks(mask_fn):
    # section centre on a lattice point (ny = nz = 33 -> 16), straight
    # daughters (angle 0): every distance below is an integer
    shape = (80, 33, 33)
    p = {"parent_radius": 6.0, "daughter_radius": 4.0, "parent_length": 30.0,
         "daughter_length": 20.0, "bifurcation_angle": 0.0}
    m = mask_fn(shape, p)
    # along the parent: distance 5 fluid, exactly 6 wall (strict <), both axes
    assert not m[10, 16, 21] and m[10, 16, 22] and not m[10, 11, 16] and m[10, 10, 16]
    # x = 0 end of the parent is inside its segment (t = 0)
    assert not m[0, 16, 21] and m[0, 16, 22]
    # straight daughter from x = 30 to x = 50, radius 4: 3 fluid, 4 wall
    assert not m[40, 16, 19] and m[40, 16, 20] and not m[40, 13, 16] and m[40, 12, 16]
    # past its end (x = 50) the distance is to the end point: x = 53, dz = 0 is
    # 3 away (fluid), x = 54 is exactly 4 (wall); (53, 16, 19) is sqrt(18) > 4
    assert not m[53, 16, 16] and m[54, 16, 16] and m[53, 16, 19]
    # y-z symmetry of a straight vessel
    assert np.array_equal(m, m[:, ::-1, :]) and np.array_equal(m, m[:, :, ::-1])
    # a daughter at 90 degrees runs along +-y from (30, 16, 16): (30, 16 + 10, 16) fluid
    p90 = dict(p, bifurcation_angle=90.0, daughter_length=12.0)
    m90 = mask_fn(shape, p90)
    assert not m90[30, 26, 16] and not m90[30, 6, 16]
    assert not m90[30, 28, 19] and m90[30, 28, 20]     # 3 and exactly 4 from the y-axis
    return m


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_closed_form_distances(impl):
    _closed_form_checks(voxelize_vessel if impl == "product" else og.vessel_mask)


@pytest.mark.parametrize("params,shape", [
    (REF, (64, 32, 32)),
    ({"parent_radius": 5.5, "daughter_radius": 3.2, "parent_length": 22.0,
      "daughter_length": 30.0, "bifurcation_angle": 41.0}, (70, 40, 24)),
    ({"parent_radius": 4.0, "daughter_radius": 4.0, "parent_length": 0.0,
      "daughter_length": 10.0, "bifurcation_angle": 10.0}, (20, 17, 15)),
])
def test_product_equals_scalar_oracle(params, shape):
    assert np.array_equal(voxelize_vessel(shape, params), og.vessel_mask(shape, params))


def test_node_accepts_the_mask():
    """The mask goes straight into LBMNode as the reference's vessel graphs
    pass it (vessel_flow_helpers.py), without a GPU: the constructor's shape
    check only."""
    from maddening_amd.node import LBMNode
    mask = voxelize_vessel((64, 32, 32), REF)
    with pytest.raises(ValueError, match="wall_mask shape"):
        LBMNode("v", 0.01, grid_shape=(64, 32, 31), wall_mask=mask, device="cpu")


@pytest.mark.gpu
def test_vessel_node_matches_oracle():
    """The vessel-flow demo's node (vessel_flow_server.py: D3Q19 64x32x32,
    Zou-He pressure inlet and outlet) on this mask: 20 LBMNode.update calls
    on the GPU (one launch each) against 20 float64 oracle steps."""
    import torch
    from maddening_amd.node import LBMNode
    from oracle import lbm_numpy as orc
    mask = voxelize_vessel((64, 32, 32), REF)
    node = LBMNode("vessel", 0.01, grid_shape=(64, 32, 32), viscosity=0.05, wall_mask=mask,
                   inlet_face="x_min", outlet_face="x_max", device="cuda")
    assert node.kernel.schedule(1) == (1, 1)
    st = node.initial_state()
    bi = {"inlet_pressure": 1.02 / 3, "outlet_pressure": 1.0 / 3}
    for _ in range(20):
        st = node.update(st, bi, 0.01)
    torch.cuda.synchronize()
    ref = orc.initial_state((64, 32, 32), 19, mask=mask.astype(np.uint8))
    out = orc.run(ref["f"], 20, tau=node.tau, mask=mask.astype(np.uint8),
                  inlet=("x_min", 1.02 / 3), outlet=("x_max", 1.0 / 3))
    v = st["velocity"].cpu().numpy()
    assert np.max(np.abs(v - out["velocity"])) / np.max(np.abs(out["velocity"])) < 1e-5
    assert np.max(np.abs(st["density"].cpu().numpy() - out["density"])) < 1e-5
    assert np.all(v[mask] == 0)
