"""CPU oracle for the BFS path (test infrastructure only; see cpu.py)."""
