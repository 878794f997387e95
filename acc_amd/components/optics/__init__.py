"""optics components."""
