"""CPU oracle of the iPregel superstep hot path (test infrastructure only)."""
