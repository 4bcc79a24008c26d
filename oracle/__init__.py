"""FP64 oracle of the Yee timestep: test infrastructure only (the checker)."""
