"""CPU oracle of the reference arithmetic. Test infrastructure only (see oracle.py)."""
