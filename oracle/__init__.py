"""FP64 CPU oracle of the two-site DMRG hot path (test infrastructure only)."""
