"""Shared inputs of the elasticity tests: tissue maps with a background and a tumor (the builders
live in glia_amd.synth, where the benchmark configurations take them too)."""
from glia_amd.synth import tissue_and_tumor, tumor_force  # noqa: F401
