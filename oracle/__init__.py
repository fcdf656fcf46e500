"""CPU oracle for Dr.Top-k -- test infrastructure only (see oracle/oracle.py)."""
