"""wp_sph -- the reference's package surface (wp_sph/__init__.py:1), backed by
the MI355X IISPH step in warpsph_amd."""
from . import solver  # noqa: F401
