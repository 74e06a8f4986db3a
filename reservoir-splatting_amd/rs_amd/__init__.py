"""Host-side mirror of the ReservoirSplatting render pass over the MI355X C ABI.

`ReservoirSplatting` keeps the reference pass's property names and meaning
(Source/RenderPasses/ReservoirSplatting/ReservoirSplatting.cpp:99-131, 182-301) and its
Python surface (`reset()`, `pixelStats`, `useFixedSeed`, `fixedSeed`; :142-156).  All
work runs in the HIP library reservoir-splatting_amd/lib/librs_amd.so; there is no
CPU fallback: without the library every entry point raises.
"""
from .abi import BUF, GATHER, TEMPORAL, BACKUP_MIS, MIS  # noqa: F401
from .pass_ import ReservoirSplatting, RSError, load_library, library_path  # noqa: F401
from .accumulate import AccumulatePass  # noqa: F401
from .replay import CameraReplay  # noqa: F401
from . import scenes  # noqa: F401
