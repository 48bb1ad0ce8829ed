"""The device pattern builder produces exactly the host builder's arrays."""
import numpy as np
import pytest
import torch

from tensormesh_amd.mesh import unit_cube_tet, unit_square_tri
from tensormesh_amd.pattern import Pattern

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["tri", "tet", "tet10", "tri6", "cfg2"])
def test_device_pattern_equals_host(kind):
    print(f"case device pattern {kind}", flush=True)
    if kind == "tri":
        m = unit_square_tri(17, 0.2, 1)
    elif kind == "tri6":
        m = unit_square_tri(9, 0.1, 1, order=2)
    elif kind == "tet":
        m = unit_cube_tet(6, 0.15, 1)
    elif kind == "tet10":
        m = unit_cube_tet(4, 0.0, 1, order=2)
    else:
        m = unit_cube_tet(64, 0.15, 0)
    conn = m.elements()
    host = Pattern.build(conn, m.n_points)
    dev = Pattern.build_device(conn.cuda(), m.n_points)
    assert dev is not None
    for name in ("row_ptr", "col_idx", "seg_ptr", "seg_slot", "slot2nnz", "inc_ptr", "inc"):
        a, b = getattr(host, name), getattr(dev, name).cpu()
        assert torch.equal(a, b), name
    assert host.nnz == dev.nnz and host.hash == dev.hash
    if conn.shape[1] <= 4:
        assert torch.equal(host.colpos(conn).cpu(), dev.colpos(conn).cpu())
        assert torch.equal(host._entry_mask.cpu(), dev._entry_mask.cpu())
