"""MI355X-native Baraff-Witkin implicit cloth step (HIP/gfx950)."""
