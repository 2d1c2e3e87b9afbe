"""CPU oracle of the star-join path (test infrastructure only; see oracle/oracle.py)."""
