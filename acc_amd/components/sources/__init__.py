"""sources components."""
