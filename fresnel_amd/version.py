"""Version of the fresnel_amd package (API level of fresnel 0.13.5)."""

version = "0.13.5+mi355x.1"
"""str: Package version."""
