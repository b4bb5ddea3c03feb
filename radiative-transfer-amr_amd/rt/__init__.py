"""Radiative-transfer problem interface (placeholder until Problem lands)."""
