"""MI355X-native PaScaL_POISSON_FFT solve path (HIP, gfx950)."""
