"""splatlight_amd: the splat frame path of Splatlight on MI355X (gfx950), behind a C ABI.

The compute path is libsplatlight_hip.so (hand-written HIP). Python here is the host mirror of the reference's
JavaScript interface and the synthetic scene generator; it computes nothing of the frame itself.
"""
from . import records, scenes  # noqa: F401

__all__ = ["records", "scenes", "native", "renderer"]
