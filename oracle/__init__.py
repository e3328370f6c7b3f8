"""CPU oracle of the rendering hot path -- test infrastructure only."""
