"""MI355X (gfx950) backend for ph-core's NArray data-parallel path.

Python access to the host-side Phase API (ph-core_amd/host, C++), which calls
the HIP kernels through the C-ABI in include/phhip.h (ph-core_amd/lib/
libphhip.so).  Importing fails if the native libraries have not been built
(`python -c "import __graft_entry__ as g; g.build()"`); there is no fallback.
"""
import os
import sys

_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib")
if _LIB not in sys.path:
    sys.path.insert(0, _LIB)

import _phase as _native  # noqa: E402
from _phase import *  # noqa: E402,F401,F403

LIB_DIR = _LIB
NATIVE_PATH = _native.__file__
