"""CPU oracle of GaussMLE.jl's per-ROI fit -- test infrastructure only."""
