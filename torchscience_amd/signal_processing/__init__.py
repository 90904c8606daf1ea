from . import filter  # noqa: F401
