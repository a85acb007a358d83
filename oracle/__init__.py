"""CPU oracle for the BFS hot path -- test infrastructure only (see oracle.py)."""
