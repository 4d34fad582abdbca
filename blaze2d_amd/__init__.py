"""blaze2d_amd: MI355X (gfx950) backend for the Blaze2D band-solve path.

Host-side mirror of the reference's interfaces for this path:

* :class:`HipBackend` / :class:`Field` - the ``SpectralBackend`` trait
  (crates/core/src/backend.rs:43-165) on device buffers.
* :class:`ThetaOperator` - ``LinearOperator`` + preconditioners
  (crates/core/src/operators/maxwell.rs, crates/core/src/preconditioners/).
* :func:`run_with_options` / :func:`run_batch` - ``bandstructure::run_core``
  (crates/core/src/drivers/bandstructure.rs:381-885) with the LOBPCG
  ``Eigensolver::solve`` (crates/core/src/eigensolver.rs:1684-2192), many band
  structures batched on one GPU.

All compute goes through libblaze2d_hip.so; there is no CPU fallback.
"""
from .bandstructure import (
    BandStructureJob,
    BandStructureResult,
    BasisAtom,
    Geometry2D,
    Lattice2D,
    RunOptions,
    dielectric,
    k_path,
    run_batch,
    run_with_options,
)
from .backend import Field, HipBackend, ThetaOperator

__all__ = [
    "BandStructureJob",
    "BandStructureResult",
    "BasisAtom",
    "Field",
    "Geometry2D",
    "HipBackend",
    "Lattice2D",
    "RunOptions",
    "ThetaOperator",
    "dielectric",
    "k_path",
    "run_batch",
    "run_with_options",
]
