"""Drop-in replacement of the reference's misc f2py module f_misc
(misc/compile_misc.py:36)."""
