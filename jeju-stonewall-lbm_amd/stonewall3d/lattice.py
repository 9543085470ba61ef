"""D3Q19 lattice tables in the reference's order (stonewall3d/lbm3d.py:20-30).

The velocity set, weights and flags are a fixed standard (D3Q19, weights 1/3, 1/18,
1/36), so they are alike by necessity; OPP / MIRZ are derived from C the same way
(opposite vector, z-mirrored vector).  MIRY (y-mirrored vector) is new: it serves the
free-slip side walls.
"""
import numpy as np

C = np.array([(0, 0, 0),
              (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
              (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0),
              (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
              (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)], dtype=np.int32)
Q = 19
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)


def _index_of(v):
    return int(np.nonzero((C == v).all(1))[0][0])


OPP = np.array([_index_of(-C[i]) for i in range(Q)], dtype=np.int32)
MIRZ = np.array([_index_of(C[i] * np.array([1, 1, -1])) for i in range(Q)], dtype=np.int32)
MIRY = np.array([_index_of(C[i] * np.array([1, -1, 1])) for i in range(Q)], dtype=np.int32)

FLUID, WALL, GROUND = 0, 1, 2


def flag_from_mask(wall):
    """Solid flags as lbm3d.py:288-290: stones WALL, the layer z=0 GROUND (overrides)."""
    flag = np.zeros(wall.shape, np.uint8)
    flag[wall] = WALL
    flag[:, :, 0] = GROUND
    return flag
