"""Bindings of the path to other array frameworks (optional dependencies)."""
