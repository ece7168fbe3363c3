"""The oracle's snapshot restatement (tools/file_io.py, src/input_output.f90:
498-728) against byte layouts built independently with the struct module:
header order and big-endian encoding, the Fortran order of the field data,
the total length counted item by item, and read/write round trips of both
formats."""
import struct

import numpy as np
import pytest

from oracle import snapshot

NX, NY, NZ, NB = 6, 4, 6, 3


def sample(nb=NB, seed=0):
    rng = np.random.default_rng(seed)
    nxg, nyg, nzg = NX + 1, NY + 1, NZ + 1
    d = {"x": rng.random(NX), "y": rng.random(NY), "z": rng.random(NZ), "xm": rng.random(NX - 1),
         "ym": rng.random(NY - 1), "zm": rng.random(NZ - 1)}
    shapes = {"U": (NX, nyg, nzg), "V": (nxg, NY, nzg), "W": (nxg, nyg, NZ), "P": (nxg, nyg, nzg),
              "Hu_interior": (NX, nyg, nzg), "Hv_interior": (nxg, NY, nzg),
              "Hw_interior": (nxg, nyg, NZ)}
    for k, s in shapes.items():
        d[k] = np.asfortranarray(rng.standard_normal(s))
    d.update(t=1.25, dt=0.01, dpdx=-3e-3, nu=1 / 2800, nxb=nb, nzb=1)
    for k, n in (("xb", nb), ("yb", nb), ("zb", nb), ("fb", 4 * nb), ("ub", 3 * nb), ("sb", nb),
                 ("normals", 3 * nb), ("tangents_1", 3 * nb), ("tangents_2", 3 * nb)):
        d[k] = rng.standard_normal(n)
    return d


@pytest.mark.parametrize("fields_only", [False, True])
def test_snapshot_bytes(tmp_path, fields_only):
    d = sample()
    p = tmp_path / "snap"
    snapshot.write(p, d, fields_only)
    raw = p.read_bytes()
    assert len(raw) == snapshot.expected_size(NX, NY, NZ, NB, fields_only)
    # mesh: int32 n then n doubles, big-endian, x first
    assert raw[:4] == struct.pack(">i", NX)
    assert raw[4:4 + 8 * NX] == struct.pack(f">{NX}d", *d["x"])
    off = sum(4 + 8 * n for n in (NX, NY, NZ, NX - 1, NY - 1, NZ - 1))
    # U header and the element U(i,j,k) at its column-major position
    assert raw[off:off + 12] == struct.pack(">3i", NX, NY + 1, NZ + 1)
    i, j, k = 4, 2, 5
    pos = off + 12 + 8 * (i + NX * (j + (NY + 1) * k))
    assert struct.unpack(">d", raw[pos:pos + 8])[0] == d["U"][i, j, k]
    if not fields_only:
        # the last item: tangents_2 with its int32 length
        tail = struct.pack(">i", 3 * NB) + struct.pack(f">{3 * NB}d", *d["tangents_2"])
        assert raw.endswith(tail)


@pytest.mark.parametrize("fields_only", [False, True])
@pytest.mark.parametrize("nb", [NB, 0])
def test_snapshot_round_trip(tmp_path, fields_only, nb):
    d = sample(nb)
    p = tmp_path / "snap"
    snapshot.write(p, d, fields_only)
    r = snapshot.read(p, fields_only)
    for k, v in r.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, d[k]) and v.shape == np.shape(d[k]), k
        else:
            assert v == d[k], k
