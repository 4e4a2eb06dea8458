"""TPB: the teapot_base map functions (reference: py/orbit/teapot_base/__init__.py)."""
import teapot_base as TPB

__all__ = ["TPB"]
