"""simkit_amd: MI355X (gfx950) implementation of simkit's implicit-FEM
backward-Euler path behind simkit's own function names and signatures.

Every numerical routine runs in the HIP library libsimkit_amd.so (built by
``python -m simkit_amd.build``) through its C ABI (include/simkit_amd.h);
PyTorch-ROCm tensors are only the device buffers. There is no CPU fallback:
without the library or a GPU, calls raise.
"""

from ._lib import SimkitAmdError  # noqa: F401
from .backtracking_line_search import backtracking_line_search  # noqa: F401
from .deformation_gradient import deformation_gradient  # noqa: F401
from .deformation_jacobian import deformation_jacobian  # noqa: F401
from .gravity_force import gravity_force  # noqa: F401
from .massmatrix import massmatrix  # noqa: F401
from .polar_svd import polar_svd  # noqa: F401
from .psd_project import psd_project  # noqa: F401
from .rotation_gradient import rotation_gradient_F  # noqa: F401
from .svd_rv import svd_rv  # noqa: F401
from .tetrahedron_volumes import tetrahedron_volumes  # noqa: F401
from .volume import volume  # noqa: F401
from .ympr_to_lame import ympr_to_lame  # noqa: F401
from . import energies, integrators, solvers  # noqa: F401

__version__ = "0.1.0"
