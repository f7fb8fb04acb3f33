"""Subpackage of pyhub_amd mirroring pyhub.core."""
