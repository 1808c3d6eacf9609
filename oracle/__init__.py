"""CPU oracle of the pseudospectral path -- test infrastructure only."""
