"""monitors components."""
