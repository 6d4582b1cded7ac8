"""Encoder surface (reference: encoder/__init__.py, encoder/implementation/stride.py)."""
from .stride import STRIDEEncoder

__all__ = ["STRIDEEncoder"]
