"""MI355X-native Hull-White one-factor trinomial-tree pricer.

Python front end of the C ABI declared in include/hwtree.h.  Import this
package through ``load()`` in the repository root modules (the directory name
contains a dot, so it is loaded by path under the name ``optionspricing_amd``).
"""
from . import formats, datasets, capi, devmem, surrogate  # noqa: F401
from .formats import Options, Yield, CALL, PUT  # noqa: F401
from .capi import price, alphas, Plan, HwtError  # noqa: F401
