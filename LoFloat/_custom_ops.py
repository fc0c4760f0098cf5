"""``LoFloat._custom_ops``: the reference module path, served by ``lofloat_amd._custom_ops`` itself (this
name is bound to that module object)."""
import sys

from lofloat_amd import _custom_ops as _impl

sys.modules[__name__] = _impl
