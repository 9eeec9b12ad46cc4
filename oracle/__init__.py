"""CPU oracle of the libirl OSEM path (test infrastructure only)."""
