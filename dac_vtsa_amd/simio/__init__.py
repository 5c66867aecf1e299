"""State hand-off files."""
