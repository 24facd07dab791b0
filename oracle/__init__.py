"""FP64 CPU oracle of the SIMP inner loop (test and baseline infrastructure only)."""
