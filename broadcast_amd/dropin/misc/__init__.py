"""Top-level ``misc`` package of the drop-in path: f_misc from
broadcast_amd/misc; other ``misc`` directories on the module path follow
(the reference's PETSc_func)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
try:
    import _path as _p
finally:
    del sys.path[0]
_p.ensure_root()
__path__ = _p.SearchPath(os.path.join(_p.PKG, "misc"), __name__)
