"""Predefined factor libraries built from the front-end ops."""
