"""MI355X-native (gfx950) DCT-preconditioned IRLS L1 phase unwrapping.

Drop-in for insardev's 2-D unwrap hot path (insardev/insardev/utils_unwrap2d.py:1082-1766):

    from core_amd import irls_unwrap_2d
    unwrapped, conncomp = irls_unwrap_2d(phase, weight, device='cuda', max_iter=10, tol=0,
                                         cg_max_iter=50, cg_tol=0)

`install()` rebinds insardev.utils_unwrap2d.irls_unwrap_2d when insardev is importable, so the
reference's Stack_unwrap2d tasks (which look the function up on that module at call time,
Stack_unwrap2d.py:12, :65-69, :124-128) run on this solver unchanged.  `core_amd.stack` has the
stack-level entry points (unwrap2d_irls / unwrap2d / unwrap2d_dataset) that send all pairs of a
burst through one batched device call (with `link_components`, the default component linking);
`core_amd.autograd.irls_unwrap_2d_diff` is the
differentiable form and `core_amd.autotune` the transform-plan tuner.
"""
from __future__ import annotations

from .linking import link_components  # noqa: F401
from .unwrap2d import (irls_unwrap_2d, irls_unwrap_2d_batch, precondition_device, resolve_device,  # noqa: F401
                       solve_device)

__version__ = "0.1.0"


def install() -> bool:
    """Rebind insardev.utils_unwrap2d.irls_unwrap_2d to this implementation (True if done)."""
    try:
        from insardev import utils_unwrap2d  # type: ignore
    except Exception:
        return False
    utils_unwrap2d.irls_unwrap_2d = irls_unwrap_2d
    return True
