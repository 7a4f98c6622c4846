"""Navigation on the device (filled in with the navigator)."""
from . import capi  # noqa: F401
