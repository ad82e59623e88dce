"""mqdmrg: host front-end of the gfx950 sigma-vector / boundary-propagation engine.

The compute path is libmqdmrg.so (HIP kernels behind the C-ABI of include/mq_sigma.h); this
package builds its inputs (symmetry indices, MPOs, MPS layouts) and drives sweeps."""
from . import capi, symmetry, models, mps  # noqa: F401

__all__ = ["capi", "symmetry", "models", "mps"]
