"""2-D geometries."""
