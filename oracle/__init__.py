"""CPU oracle of the scatterer engine (test infrastructure)."""
