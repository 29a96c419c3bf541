"""Helpers mirroring GPUSimulators/helpers (initial conditions)."""
