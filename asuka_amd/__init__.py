"""MI355X-native integral-direct four-center J/K engine (HIP kernels behind a C-ABI)."""

__version__ = "0.1.0"
