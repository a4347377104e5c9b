"""Synthetic SVGF inputs for the tests (the producer lives in lighthouse2_amd.scenes)."""
from lighthouse2_amd.scenes import gbuffer, hdr_to_rgb32, pack_normal2  # noqa: F401
