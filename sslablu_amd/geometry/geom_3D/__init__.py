"""3-D geometries."""
