"""CPU oracle of the index + query path (test infrastructure only)."""
