"""CPU oracle of the path-tracing core: test infrastructure only (see oracle_pt.h)."""
