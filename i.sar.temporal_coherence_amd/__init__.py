"""MI355X (gfx950) backend of i.sar.temporal_coherence's phase-linking path.

The product is libtcoh_hip.so (HIP kernels behind the C ABI of
include/tcoh_hip.h, a drop-in for the reference's ocl.c). This package holds
its sources (csrc/, Makefile) and the Python binding (api).

The directory name is not a Python identifier; load it by path, e.g.

    spec = importlib.util.spec_from_file_location(
        "tcoh_amd", path + "/__init__.py", submodule_search_locations=[path])
"""

from . import api  # noqa: F401
from .api import (Backend, Batch, make_settings, pair_list, prepare_strip,  # noqa: F401
                  temporal_coherence)
