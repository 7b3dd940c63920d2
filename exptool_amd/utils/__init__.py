"""Table I/O, orbit integration and samplers."""
