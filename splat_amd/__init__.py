"""MI355X-native per-frame tracing hot path of SPLAT (gfx950 HIP kernels behind a C ABI)."""
