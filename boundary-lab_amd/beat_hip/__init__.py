"""beat_hip: MI355X-native exterior Helmholtz BEM solver (drop-in for Boundary Lab's `beat_rocm`)."""

__version__ = "0.1.0"
