"""CPU oracle of the py-lddmm hot path (test infrastructure only)."""
