"""Particle containers and readers."""
