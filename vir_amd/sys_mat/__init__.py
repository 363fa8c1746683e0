"""Projector families: siddon (ray-driven Siddon), dd (distance-driven)."""
__all__ = ["siddon"]
