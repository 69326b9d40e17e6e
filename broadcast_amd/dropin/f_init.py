"""Top-level ``f_init`` of the drop-in path (compilef90.py:17)."""
import os
import sys

_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if _ROOT not in sys.path:
    sys.path.append(_ROOT)

from broadcast_amd.f_init import set_bndbl_2d  # noqa: E402,F401
