"""MI355X SPH/PCISPH step library (package placeholder)."""
