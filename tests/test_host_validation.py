"""Host-side argument validation of the C ABI (no GPU: rejected before any
device call), mirroring the reference's stops on invalid sizes
(src/initialization.f90:46-48)."""
import os

import pytest

from channel3d_parallel_amd import C3DError, Channel, _lib

needs_lib = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libc3d.so not built")


@needs_lib
@pytest.mark.parametrize("nx,ny,nz,gt", [
    (35, 65, 34, 1),    # odd nx
    (34, 65, 35, 1),    # odd nz
    (4, 65, 34, 1),     # too small
    (34, 3, 34, 1),     # too few y faces
    (2 + 80, 65, 34, 1),  # FFT length 80 unsupported
    (34, 65, 34, 7),    # unknown grid type
])
def test_invalid_sizes_rejected(nx, ny, nz, gt):
    with pytest.raises(C3DError) as e:
        Channel(nx, ny, nz, grid_type=gt)
    assert e.value.code == -1


@needs_lib
def test_unknown_precision_code_is_rejected():
    with pytest.raises(C3DError):
        Channel(34, 65, 34, precision=7)
