"""CPU oracle of the CG-SENSE path (test infrastructure only; see ora.py)."""
