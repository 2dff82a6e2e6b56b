"""CPU oracle of the Census-SGM matcher -- test infrastructure only (see oracle.hpp)."""
