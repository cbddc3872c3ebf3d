"""CPU oracle for the FBP hot path (test infrastructure only)."""
