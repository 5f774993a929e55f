"""MI355X LeCo_fix codec: host-side mirror of Codecset::Leco_int<T> over the
gfx950 HIP library libleco_fix.so (C ABI in include/leco_fix.h)."""
from ._lib import LECO_FIX_TAIL_PAD, BlockParam, LecoError, lib  # noqa: F401
from .codec import Leco_int  # noqa: F401


def __getattr__(name):
    # Column needs torch; import it lazily so the host-buffer surface works without it.
    if name == "Column":
        from .column import Column
        return Column
    raise AttributeError(name)
