"""Component plugin layer (acc/components/)."""
