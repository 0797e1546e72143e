"""Correctly spelt alias of :mod:`fan_equiangluar` (the reference's numpy module
name is misspelt; its cupy module is spelt fan_equiangular)."""
from .fan_equiangluar import *  # noqa: F401,F403
from .fan_equiangluar import ramp_filter, fbp_bp, distance_driven_fp, distance_driven_bp  # noqa: F401
