"""MI355X-native augmented-Lagrangian low-rank SDP search of lagbound (HIP kernels behind a C ABI)."""
__version__ = "0.1.0"
