"""The reference's f2py module names, resolvable without editing its drivers.

The reference imports its compiled modules as ``srcfv.f_geom``,
``srcfv.f_bnd``, ``srcfv.f_sch``, ``srcfv.f_lin``, ``srcfv.f_adj``,
``srcfv.f_norm``, ``srcfv.f_dz``, ``misc.f_misc`` and ``f_init``
(BROADCAST_npz.py:15-34, cylinder.py:16-29).  Putting this directory first on
the module search path, ahead of the reference's own tree, makes those names
load the MI355X implementations of ``broadcast_amd.srcfv``,
``broadcast_amd.misc`` and ``broadcast_amd.f_init``:

    PYTHONPATH=<repo>/broadcast_amd/dropin python BROADCAST_npz.py

The ``srcfv`` and ``misc`` packages here extend their search path with the
same-named directories found later on the module path (pkgutil.extend_path),
so modules this project does not provide -- ``srcfv.f_lhs``,
``misc.PETSc_func`` -- still load from the reference's tree.
"""
import os
import sys

PATH = os.path.dirname(os.path.abspath(__file__))


def install():
    """Put PATH first on sys.path (the PYTHONPATH line above, in-process)."""
    if PATH not in sys.path:
        sys.path.insert(0, PATH)
