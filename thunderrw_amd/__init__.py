"""thunderrw_amd: MI355X-native Gather-Move-Update random-walk step engine.

The compute path is libthunderrw.so (HIP kernels for gfx950 behind the C ABI
of include/thunderrw.h).  Importing the package does not load it; the first
call does, and raises if it has not been built.
"""
from ._abi import TrwError, exported_names, lib  # noqa: F401
from .engine import (ALIAS_DTYPE, APP_DEFAULTS, Context, DeviceBuffer, DeviceGraph, DeviceTimer,  # noqa: F401
                     HostGraph, Server, device_count, make_io, make_params, replay_fill, sources_all, sources_random,
                     sources_single, synchronize, walk, walk_async, walk_mix_async, walk_noalloc)

__all__ = ["TrwError", "Context", "DeviceBuffer", "DeviceGraph", "DeviceTimer", "HostGraph", "Server", "device_count",
           "make_io", "make_params", "replay_fill", "sources_all", "sources_random", "sources_single", "synchronize",
           "walk", "walk_async", "walk_mix_async", "walk_noalloc", "ALIAS_DTYPE", "APP_DEFAULTS", "exported_names", "lib"]
