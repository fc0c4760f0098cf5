"""``LoFloat.formats``: the reference module path, served by ``lofloat_amd.formats`` itself (this
name is bound to that module object)."""
import sys

from lofloat_amd import formats as _impl

sys.modules[__name__] = _impl
