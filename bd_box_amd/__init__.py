"""bd_box_amd: MI355X (gfx950) backend for BD_BOX's Ermak-McCammon step with
hydrodynamic interactions.

The compute path is the HIP library libbdbox_hip.so behind the C ABI of
include/bdbox_hip.h; `bd_box_amd.system.BDSystem` mirrors the reference's hot-path
interface on top of it.  `bd_box_amd.workloads` writes and reads the reference's
`.prm` / `.str` formats and builds the seeded benchmark systems.
"""
from . import workloads  # noqa: F401

__all__ = ["workloads", "system"]
