"""gales_mc: MI355X-native fluid_mc element assembly (Python host binding).

The compute path is libgales_mc.so (hand-written HIP for gfx950) reached
through its C-ABI (include/gales_mc.h); this package holds the ctypes binding,
the synthetic benchmark cases and the torch.autograd adjoint binding.
"""
from . import abi, cases  # noqa: F401
from .lib import Context, GmcError, load  # noqa: F401
