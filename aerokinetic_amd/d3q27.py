"""D3Q27 velocity set in the package's direction order.

Direction i = 9 (cx+1) + 3 (cy+1) + (cz+1), so opposite(i) = 26 - i and the rest
population is i = 13.  Weights 8/27, 2/27, 1/54, 1/216 for |c|^2 = 0, 1, 2, 3.
The order is internal; the reference derives opposites by lookup and tests that
results do not depend on the order (tests/test_refinement.py:153-160).
"""
import numpy as np

Q = 27
C = np.array([(i // 9 - 1, (i // 3) % 3 - 1, i % 3 - 1) for i in range(Q)], dtype=np.int64)
CX, CY, CZ = C[:, 0], C[:, 1], C[:, 2]
OPPOSITE = np.arange(Q)[::-1].copy()
W = np.array([{0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}[int(n)] for n in (C * C).sum(axis=1)])
REST = 13
