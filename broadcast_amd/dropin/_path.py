"""Search path of the drop-in ``srcfv`` / ``misc`` packages."""
import os
import pkgutil
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # .../broadcast_amd
ROOT = os.path.dirname(PKG)


def ensure_root():
    """broadcast_amd itself must be importable by the modules these names load."""
    if ROOT not in sys.path:
        sys.path.append(ROOT)


class SearchPath(list):
    """This project's directory first, then every other same-named package
    directory on the module path, recomputed at each lookup: the driver
    script's own directory (the reference's tree) joins sys.path only after
    these packages were imported."""

    def __init__(self, own, name):
        super().__init__([own])
        self._own, self._name = own, name

    def __iter__(self):
        return iter(pkgutil.extend_path([self._own], self._name))
