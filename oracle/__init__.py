"""CPU oracle of the particle hot path (test infrastructure only)."""
