"""Spherical (SL) and cylindrical (EOF) bases, field glue and table generation."""
