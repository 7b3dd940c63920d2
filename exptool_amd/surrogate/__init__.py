"""Learned surrogate of a frozen BFE field (no counterpart in the reference)."""
