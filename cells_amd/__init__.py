"""cells_amd: the MI355X-native batched tissue step of Eidosoma/Cells.

Hand-written HIP kernels for gfx950 behind a C-ABI (include/cells_hip.h, libcells_hip.so); this
package is the host side: world packing (world), symbolic analysis (symbolic), device ensembles
(ensemble) and the reference's model contract (model).
"""

from cells_amd import world  # noqa: F401

__all__ = ["world", "symbolic", "ensemble", "model", "lib", "build"]
