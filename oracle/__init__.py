"""CPU oracle of the lbm3d step (test infrastructure only)."""
