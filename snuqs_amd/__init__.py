"""snuqs_amd -- MI355X-native state-vector engine for SnuQS.

``snuqs_amd._C`` mirrors the reference's native module ``braket.snuqs._C``
(src/braket/snuqs/csrc/pymodule.cc:66-210) over the C ABI of
``libsnuqs_hip.so`` (include/snuqs_hip.h).  There is no CPU execution path:
importing the package fails if the HIP library has not been built.
"""
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

try:
    from . import _C  # noqa: F401
except ImportError as e:  # fail loudly: no fallback path exists
    raise ImportError(
        "snuqs_amd._C is not built (or libsnuqs_hip.so is missing); run "
        "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root"
    ) from e

from . import autotune as _autotune  # noqa: E402

_autotune.load()  # launch-shape knobs measured on the target GPU, if present

__all__ = ["_C"]
