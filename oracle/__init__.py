"""FP64 CPU oracle of the band path -- test infrastructure only (see oracle.py)."""
