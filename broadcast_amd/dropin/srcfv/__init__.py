"""Top-level ``srcfv`` package of the drop-in path: its modules are those of
broadcast_amd/srcfv (f_sch, f_lin, f_adj, f_bnd, f_geom, f_norm, f_dz); other
``srcfv`` directories on the module path follow, for the modules this
project does not provide (the reference's f_lhs)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
try:
    import _path as _p
finally:
    del sys.path[0]
_p.ensure_root()
__path__ = _p.SearchPath(os.path.join(_p.PKG, "srcfv"), __name__)
