"""Shared inputs of the mass-effect forward tests (built by glia_amd.synth.mass_effect_case: tissue
with all four classes, a tumor seed and a background that closes the partition of unity)."""
from glia_amd.synth import mass_effect_case  # noqa: F401
