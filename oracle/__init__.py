"""FP64 CPU oracle of the OpenTM homogenization path (test infrastructure only)."""
