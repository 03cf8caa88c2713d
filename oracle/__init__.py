"""FP64 CPU oracle for the particle step (test infrastructure only)."""
