"""Which lbm_step calls run two steps per launch (LBM_FLAG_PAIR_*), checked on
the CPU through lbm_step_schedule, which makes no HIP call.

A D3Q19 pull sequence of n steps is n + 1 units (a collision, n - 1 fused
steps, a stream + moments); with two-step launches every launch but at most
one carries two units, so the call makes (n + 2) // 2 launches, (n + 1) // 2
of them two-step launches, and a single update is one launch."""
import ctypes

import pytest


def _desc(dims, storage=4, bb=0, flags=0, faces=False, q=19, force=None):
    from maddening_amd import _lib
    d = _lib.LbmDesc()
    d.q, d.storage, d.bb_mode, d.tau, d.flags = q, storage, bb, 0.8, flags
    for i, n in enumerate(dims):
        d.dims[i] = n
    if faces:
        d.inlet_face, d.inlet_rho = 5, 1.0
    for i, v in enumerate(force or ()):
        d.force[i] = v
    return d


def _schedule(d, nsteps, with_fields=0):
    from maddening_amd import _lib
    nl, npair = ctypes.c_int(-1), ctypes.c_int(-1)
    _lib.check(_lib.lib().lbm_step_schedule(ctypes.byref(d), nsteps, with_fields,
                                            ctypes.byref(nl), ctypes.byref(npair)),
               "lbm_step_schedule")
    return nl.value, npair.value


def _folded(nsteps):
    return (nsteps + 2) // 2, (nsteps + 1) // 2


@pytest.mark.parametrize("storage", [0, 4])  # f32, f32dev
@pytest.mark.parametrize("nsteps", [1, 2, 3, 4, 20, 1000, 1001])
def test_auto_pairs_at_256_cubed(storage, nsteps):
    from maddening_amd import _lib
    d = _desc((256, 256, 256), storage, flags=_lib.LBM_FLAG_TRI_OFF)
    assert _schedule(d, nsteps) == _folded(nsteps)
    if nsteps % 2 or nsteps < 4:
        assert _schedule(_desc((256, 256, 256), storage), nsteps) == _folded(nsteps)


def _tri(nsteps):
    # (C, SC, SC), (nsteps - 4) / 2 pairs of SC, (SC, S)
    return nsteps // 2, nsteps // 2


@pytest.mark.parametrize("nsteps", [4, 6, 20, 1000])
def test_even_sequences_start_with_three_units(nsteps):
    """An even nsteps >= 4 has an odd number of units; with 32- or 16-bit
    storage, periodic, no force, rows of whole 64-column tiles, the first
    launch runs three units instead of (C, SC) plus a single SC."""
    from maddening_amd import _lib
    for dims in [(256, 256, 256), (128, 128, 512), (64, 256, 128), (512, 256, 128)]:
        assert _schedule(_desc(dims), nsteps) == _tri(nsteps), dims
    for storage in (0, 2, 3):  # f32, f16dev, bf16dev
        for dims in [(256, 256, 256), (64, 256, 128)]:
            assert _schedule(_desc(dims, storage=storage), nsteps) == _tri(nsteps), dims
    # forced two-step launches on a small lattice take it too
    assert _schedule(_desc((3, 5, 64), flags=_lib.LBM_FLAG_PAIR_FORCE), nsteps) == _tri(nsteps)
    # not: a force, fields, the reference wall rule, rows that are not whole
    # tiles, PAIR_OFF, TRI_OFF
    for d in [_desc((256, 256, 256), storage=0, flags=_lib.LBM_FLAG_TRI_OFF),
              _desc((256, 256, 256), storage=2, flags=_lib.LBM_FLAG_TRI_OFF),
              _desc((256, 256, 256), force=(0.0, 1e-5, 0.0)),
              _desc((256, 256, 256), bb=1), _desc((160, 160, 160)),
              _desc((256, 256, 256), flags=_lib.LBM_FLAG_TRI_OFF)]:
        assert _schedule(d, nsteps) == _folded(nsteps)
    assert _schedule(_desc((256, 256, 256)), nsteps, with_fields=1) == _folded(nsteps)
    off = _schedule(_desc((256, 256, 256), flags=_lib.LBM_FLAG_PAIR_OFF), nsteps)
    assert off == (nsteps + 1, 0)


# lattices where two-step launches measured faster than the best one-step
# launch shape: at least 2^21 cells, a and b at least 64 cells, whole rows
# of 64 to 256 cells in multiples of 32, or rows of more than 341 cells in
# 64-column tiles that cover the row with at most 10 % of extra columns
@pytest.mark.parametrize("dims", [(128, 128, 128), (64, 256, 128), (256, 128, 64),
                                  (160, 160, 160), (192, 192, 192), (256, 256, 160),
                                  (320, 320, 192), (512, 256, 128), (64, 512, 64),
                                  (128, 128, 512), (256, 256, 512), (64, 64, 2048),
                                  (96, 96, 1000), (64, 96, 352), (64, 96, 700)])
def test_auto_pairs_on_measured_winning_shapes(dims):
    assert _schedule(_desc(dims), 9) == (5, 5)


def test_auto_keeps_single_launches_where_pairs_measured_slower():
    from maddening_amd import _lib
    # 64-bit storage; fewer than 2^21 cells (96^3 and 100^3 measured slower
    # with two-step launches); rows not a multiple of 32 (200x200x100 broke
    # even), shorter than 64 or longer than 256 cells; a or b thinner than
    # 64 cells; halfway bounce-back, the push scheme, PAIR_OFF
    cases = [_desc((256, 256, 256), storage=1), _desc((96, 96, 96)), _desc((100, 100, 100)),
             _desc((127, 128, 128)), _desc((16, 16, 256)), _desc((200, 200, 100)),
             _desc((256, 256, 100)), _desc((1024, 64, 32)), _desc((256, 128, 288)),
             _desc((256, 256, 320)), _desc((32768, 1, 64)), _desc((2, 4096, 256)),
             _desc((64, 64, 500)), _desc((32, 128, 1024)), _desc((128, 16, 1024)),
             _desc((63, 256, 256)), _desc((256, 63, 256)), _desc((256, 256, 256), bb=2),
             _desc((64, 96, 342)), _desc((64, 96, 400)),   # tiles 12 % / 12 % over the row
             _desc((256, 256, 256), flags=_lib.LBM_FLAG_PAIR_OFF)]
    for d in cases:
        assert _schedule(d, 100) == (101, 0)
    d = _desc((256, 256, 256))
    d.scheme = _lib.LBM_SCHEME_PUSH
    assert _schedule(d, 100) == (101, 0)


def test_force_flag_reaches_every_row_length():
    from maddening_amd import _lib
    F = _lib.LBM_FLAG_PAIR_FORCE
    # whole rows up to 341 cells, 64-column tiles beyond
    for dims in [(1, 1, 64), (33, 50, 64), (128, 128, 128), (3, 2, 256), (8, 8, 96),
                 (2, 3, 341), (1, 1, 1), (8, 8, 342), (8, 8, 512), (2, 1, 2000)]:
        assert _schedule(_desc(dims, flags=F), 9) == (5, 5)
    # the general form: reference walls, fields, pressure faces
    for d in [_desc((8, 8, 64), bb=1, flags=F), _desc((8, 8, 64), faces=True, flags=F)]:
        assert _schedule(d, 9) == (5, 5)
    assert _schedule(_desc((8, 8, 64), flags=F), 9, with_fields=1) == (5, 5)
    # never under halfway bounce-back
    assert _schedule(_desc((8, 8, 64), bb=2, flags=F), 9) == (10, 0)
    # 16-bit storage in every form when forced; never 64-bit
    for d in [_desc((8, 8, 64), storage=2, flags=F), _desc((8, 8, 512), storage=3, flags=F),
              _desc((8, 8, 64), storage=2, bb=1, flags=F),
              _desc((8, 8, 64), storage=3, faces=True, flags=F)]:
        assert _schedule(d, 9) == (5, 5)
    assert _schedule(_desc((8, 8, 64), storage=1, flags=F), 9) == (10, 0)


def test_16_bit_auto_pairs_only_in_the_plain_form():
    """16-bit deviation storage takes two-step launches by itself only where
    the march runs its plain form (periodic, no force, no fields, no
    faces) on a shape of the rule; its other forms spill registers."""
    for storage in (2, 3):  # f16dev, bf16dev
        for dims in [(256, 256, 256), (512, 256, 128), (256, 256, 160), (128, 128, 1024)]:
            assert _schedule(_desc(dims, storage=storage), 9) == (5, 5)
        for d in [_desc((256, 256, 256), storage=storage, bb=1),
                  _desc((256, 256, 256), storage=storage, faces=True),
                  _desc((256, 256, 256), storage=storage, force=(0.0, 0.0, 1e-5)),
                  _desc((96, 96, 96), storage=storage)]:
            assert _schedule(d, 9) == (10, 0)
        assert _schedule(_desc((256, 256, 256), storage=storage), 9, with_fields=1) == (10, 0)


def test_phase_3_rejected_where_pairs_do_not_apply():
    from maddening_amd import _lib
    L = _lib.lib()
    d = _desc((8, 8, 64), storage=1, flags=_lib.LBM_FLAG_PAIR_FORCE)
    rc = L.lbm_pull_phase(None, ctypes.byref(d), 3, ctypes.c_void_p(4096), ctypes.c_void_p(8192),
                          None, None, None, None, None, None, None)
    assert rc == -2 and b"two steps per launch" in L.lbm_last_error()
    for phase in (3, 4, 5):   # sequences: 32- and 16-bit storage only
        assert L.lbm_pull_phase(None, ctypes.byref(d), phase, ctypes.c_void_p(4096),
                                ctypes.c_void_p(8192), None, None, None, None, None, None,
                                None) == -2
    d64 = _desc((8, 8, 64), storage=1)
    assert L.lbm_pull_phase(None, ctypes.byref(d64), 6, ctypes.c_void_p(4096),
                            ctypes.c_void_p(8192), None, None, None, None, None, None, None) == -2
    dh = _desc((8, 8, 64), bb=2)
    assert L.lbm_pull_phase(None, ctypes.byref(dh), 6, ctypes.c_void_p(4096),
                            ctypes.c_void_p(8192), ctypes.c_void_p(64), None, None, None, None,
                            None, None) == -2
    # phase 7 (C, SC, SC): 32- or 16-bit storage, periodic, no force, no
    # fields, rows of at least 64 cells
    for dd in [d, _desc((8, 8, 64), storage=1),
               _desc((8, 8, 32)), _desc((8, 8, 64), force=(1e-5, 0.0, 0.0))]:
        assert L.lbm_pull_phase(None, ctypes.byref(dd), 7, ctypes.c_void_p(4096),
                                ctypes.c_void_p(8192), None, None, None, None, None, None,
                                None) == -2
        assert b"three units per launch" in L.lbm_last_error()
    assert L.lbm_pull_phase(None, ctypes.byref(d), 8, ctypes.c_void_p(4096),
                            ctypes.c_void_p(8192), None, None, None, None, None, None, None) == -1


def test_d2q9_pairs_from_2_to_the_21_cells():
    from maddening_amd import _lib
    for dims in [(2048, 2048), (8192, 8192), (2048, 8192), (1536, 1536), (1024, 2048)]:
        assert _schedule(_desc(dims, q=9), 200) == (2 + 99 + 1, 99)
    for d in [_desc((1024, 1024), q=9), _desc((8192, 8192), q=9, storage=2),
              _desc((8192, 8192), q=9, bb=2), _desc((8192, 8192), q=9, flags=_lib.LBM_FLAG_PAIR_OFF)]:
        assert _schedule(d, 200) == (201, 0)
    assert _schedule(_desc((2048, 2048), q=9), 200, with_fields=1) == (201, 0)
    F = _lib.LBM_FLAG_PAIR_FORCE
    for dims in [(1, 1), (7, 64), (33, 961)]:
        assert _schedule(_desc(dims, q=9, flags=F), 9) == (6, 4)


def test_d2q9_16_bit_pairs_only_when_forced():
    from maddening_amd import _lib
    F = _lib.LBM_FLAG_PAIR_FORCE
    for storage in (2, 3):  # f16dev, bf16dev
        assert _schedule(_desc((8192, 8192), storage=storage, q=9), 200) == (201, 0)
        assert _schedule(_desc((64, 100), storage=storage, q=9, flags=F), 9) == (6, 4)
    # never float64
    assert _schedule(_desc((64, 100), storage=1, q=9, flags=F), 9) == (10, 0)


def test_single_update_is_one_launch():
    """nsteps = 1 (LBMNode.update's call): one two-step launch that collides
    the stored state and writes the streamed state with its moments, for
    D3Q19 with 32- or 16-bit storage, periodic or the reference wall rule,
    with or without fields and faces, any lattice shape, and no workspace.
    Float64, halfway bounce-back, D2Q9 and PAIR_OFF keep the push launch plus
    the moments pass."""
    from maddening_amd import _lib
    L = _lib.lib()
    for storage in (0, 2, 3, 4):
        for d in [_desc((256, 256, 256), storage), _desc((64, 32, 32), storage, bb=1),
                  _desc((5, 3, 7), storage, faces=True), _desc((1, 1, 1), storage),
                  _desc((8, 8, 1000), storage, bb=1)]:
            assert _schedule(d, 1) == (1, 1)
            assert _schedule(d, 1, with_fields=1) == (1, 1)
            assert L.lbm_workspace_size(ctypes.byref(d), 1) == 0
    for d in [_desc((64, 32, 32), storage=1), _desc((64, 32, 32), bb=2),
              _desc((64, 32), q=9), _desc((64, 32, 32), flags=_lib.LBM_FLAG_PAIR_OFF)]:
        assert _schedule(d, 1) == (2, 0)
    # an explicit push scheme stays push
    d = _desc((64, 32, 32))
    d.scheme = _lib.LBM_SCHEME_PUSH
    assert _schedule(d, 1) == (2, 0)
    # pull with PAIR_OFF: collide, then stream + moments
    d = _desc((64, 32, 32), flags=_lib.LBM_FLAG_PAIR_OFF)
    d.scheme = _lib.LBM_SCHEME_PULL
    assert _schedule(d, 1) == (2, 0)


@pytest.mark.parametrize("dims", [(128, 128, 256), (256, 256, 256)])
def test_auto_general_pairs_with_walls_fields(dims):
    """The reference wall rule and force or tau fields take the march's
    general form under the same rule (the 128x128x256 sphere channel: 0.083
    against 0.115 ms per step). Pressure faces keep one step per launch in
    sequences (their form measured slower), but a single update still takes
    the one-launch form."""
    for d, fields in [(_desc(dims, bb=1), 0), (_desc(dims), 1), (_desc(dims, bb=1), 1)]:
        assert _schedule(d, 500, with_fields=fields) == _folded(500)
    for d, fields in [(_desc(dims, faces=True), 0), (_desc(dims, bb=1, faces=True), 1)]:
        assert _schedule(d, 500, with_fields=fields) == (501, 0)
        assert _schedule(d, 1, with_fields=fields) == (1, 1)
