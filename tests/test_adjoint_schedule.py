"""Checkpoint schedule of the differentiable rollout (maddening_amd.adjoint),
checked on the CPU."""


def test_rollout_auto_checkpoint_choice():
    """checkpoint="auto" keeps every state only when they fit in the stated
    share of free memory (CPU: the choice alone, with free memory given)."""
    from maddening_amd.adjoint import AUTO_MEMORY_SHARE, resolve_checkpoint
    state = 128 * 128 * 256 * 19 * 4          # config-4 state, float32: 318.8 MB
    free = 280 * 2 ** 30                        # an idle 288 GB MI355X
    assert resolve_checkpoint(state, 500, "auto", free_bytes=free) == "all"
    assert resolve_checkpoint(state, 5000, "auto", free_bytes=free) == "sqrt"
    edge = int(AUTO_MEMORY_SHARE * free) // state
    assert resolve_checkpoint(state, edge, "auto", free_bytes=free) == "all"
    assert resolve_checkpoint(state, edge + 1, "auto", free_bytes=free) == "sqrt"
    for other in ("sqrt", "all", 7):
        assert resolve_checkpoint(state, 500, other, free_bytes=0) == other
