"""CPU checkers for the HIP backend (test infrastructure only; see pyoracle.py)."""
