"""paratreet_amd: ParaTreeT's gravity hot path on MI355X (gfx950).

The compute path is libpt_gravity.so (hand-written HIP kernels behind the C ABI
of include/pt_gravity.h); this package holds its Python binding
(``gravity.Driver`` mirroring the reference's Configuration / GravityVisitor /
TreeSpec / Driver interface) and seeded initial-condition generators.
"""
from . import ics  # noqa: F401

__all__ = ["ics", "gravity"]
