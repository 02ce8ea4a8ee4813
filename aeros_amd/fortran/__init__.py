"""Fortran drop-in for the reference's aeros_spectral module (iso_c_binding over libaer_sht.so)."""
