"""torpedo_amd: MI355X-native Gaussian rasterization (HIP kernels behind a C ABI)."""
