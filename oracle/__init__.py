"""FP64 CPU oracle of the RRTMGP radiation path (test infrastructure only)."""
