"""sol::Geometry on the host (CPU only): a TriangleMesh answers intersection queries on the GPU once
attached to a context; before that, a query fails with an error instead of falling back to a CPU
traversal."""
import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F


def test_unattached_mesh_refuses_queries():
    s = sol_amd.Scene.cornell_box()
    rays = sol_amd.gen_rays(0, 4)
    for any_hit in (False, True):
        for batched in (False, True):
            with pytest.raises(F.SolError, match="not attached to a GPU context"):
                s.geometry_intersect(rays, any_hit=any_hit, batched=batched)


def test_empty_query_is_a_no_op():
    s = sol_amd.Scene.cornell_box()
    empty = np.zeros((0, 8), np.float32)
    assert s.geometry_intersect(empty, batched=True)["hit"].size == 0
