"""``LoFloat.utils``: the reference module path, served by ``lofloat_amd.utils`` itself (this
name is bound to that module object)."""
import sys

from lofloat_amd import utils as _impl

sys.modules[__name__] = _impl
