"""EnsembleMI355X host package: the reference's run_dynamics / Simulation API backed by
libnqcd_mi355x.so (hand-written gfx950 kernels behind a C ABI)."""
from . import _lib  # noqa: F401
