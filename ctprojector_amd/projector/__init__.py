"""Projector front ends: ``numpy`` (host arrays) and ``cupy`` (device tensors)."""
