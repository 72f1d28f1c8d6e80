"""Kernel tables under the reference's names (wp_sph/solver/sph_funcs.py:43-118).

``W_table`` / ``grad_W_table`` are float32 numpy arrays for the reference's
default smoothing length; ``spline_W`` is the host float32 lookup. The solver
uploads its own tables (sized by its config) to the device.
"""
from warpsph_amd.tables import GRAD_W_SIZE, W_SIZE, make_tables, spline_w_lookup

from ..main import SMOOTHING_LENGTH

W_table = None
grad_W_table = None


def precompute_tables():
    """Create the W and grad-W/r tables (sph_funcs.py:51-86)."""
    global W_table, grad_W_table
    W_table, grad_W_table = make_tables(SMOOTHING_LENGTH, W_SIZE, GRAD_W_SIZE)


precompute_tables()


def spline_W(r, table=None):
    """Table lookup of W(r) (sph_funcs.py:92-102), float32."""
    return spline_w_lookup(r, W_table if table is None else table, SMOOTHING_LENGTH)
