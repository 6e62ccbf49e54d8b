"""Per-simplex measure (volume.py:14-41); tetrahedra only on this path."""

import numpy as np

from .tetrahedron_volumes import tetrahedron_volumes


def volume(V, F):
    """(m, 1) tet volumes. Segment and triangle meshes are outside this package's scope."""
    if np.asarray(F).shape[1] != 4:
        raise NotImplementedError("simkit_amd.volume supports tetrahedra (F.shape[1] == 4) only")
    return tetrahedron_volumes(V, F).reshape(-1, 1)
