"""``LoFloat.layers``: the reference module path, served by ``lofloat_amd.layers`` itself (this
name is bound to that module object)."""
import sys

from lofloat_amd import layers as _impl

sys.modules[__name__] = _impl
