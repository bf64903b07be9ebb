"""The Dirac gamma matrices as printed in the reference's documentation
(doc/qla.tex:1063-1113, chiral basis, gamma_1..gamma_4 = directions mu=0..3),
typed independently of the generator's tables; gamma_5 = g1 g2 g3 g4."""
import numpy as np

i = 1j
G = np.array([
    [[0, 0, 0, i], [0, 0, i, 0], [0, -i, 0, 0], [-i, 0, 0, 0]],
    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
    [[0, 0, i, 0], [0, 0, 0, -i], [-i, 0, 0, 0], [0, i, 0, 0]],
    [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]],
], dtype=complex)
G5 = G[0] @ G[1] @ G[2] @ G[3]
