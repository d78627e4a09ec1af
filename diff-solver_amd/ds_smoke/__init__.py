"""MI355X (gfx950) smoke solver: forward step, hand-written VJP and Jacobi-PCG projection of
Xayah-Graphics/diff-solver's smoke model, behind a C ABI (include/ds/ds_smoke.h)."""
from . import abi, scenes  # noqa: F401

__all__ = ["abi", "scenes", "solver"]
