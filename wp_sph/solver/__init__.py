from .IISPH import IISPH  # noqa: F401
