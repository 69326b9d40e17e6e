"""Drop-in replacements of the reference's srcfv f2py modules
(srcfv/compile_*.py): f_sch, f_lin, f_bnd, f_geom, f_norm, with the same
routine names and positional call forms, computing on the MI355X."""
