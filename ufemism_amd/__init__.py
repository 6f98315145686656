"""MI355X-native UFEMISM ice-dynamics hot path (SIA, SSA Picard + five-colour
SOR, neighbour-function operators, mass continuity) behind a C-ABI."""
