"""Binds the reference's top-level module names (srcfv, misc, f_init) to this
project while the interpreter starts, before the driver script's own
directory -- the reference's tree, whose srcfv/ and misc/ would otherwise win
-- joins the module path; then runs the interpreter's own sitecustomize."""
import importlib.machinery
import importlib.util
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))


def _bind():
    import srcfv  # noqa: F401
    import misc  # noqa: F401
    import f_init  # noqa: F401


def _chain():
    rest = [p for p in sys.path if os.path.abspath(p or os.curdir) != _HERE]
    spec = importlib.machinery.PathFinder.find_spec("sitecustomize", rest)
    if spec is not None and spec.loader is not None:
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)


_bind()
_chain()
