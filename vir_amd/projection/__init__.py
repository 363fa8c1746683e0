"""Stored-system-matrix projection operators (reference projection/)."""
__all__ = ["proj"]
