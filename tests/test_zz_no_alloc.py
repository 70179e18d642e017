"""tsg_field_o1_noalloc (staging in the library's static device array)
against tsg_field_o1 through the C ABI: two calls in a row on one grid, the
second with fewer nodes, both bit-exact; a count above the staging limit is
refused with an error."""
import sys

import numpy as np
import pytest

import tapsim_modified_amd as tsg
from tests.conftest import FIXTURE_SPECIES

pytestmark = pytest.mark.gpu


def log(msg):
    print("[gpu-noalloc] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", [tsg.PLAIN_F64, tsg.FAITHFUL_F32])
def test_field_o1_noalloc_matches_field_o1(tip_small, mode):
    log("no-alloc field_o1 mode %d" % mode)
    mesh = tsg.Mesh(tip_small["xyz"], tip_small["tets"], tip_small["ref_neigh"])
    g = tsg.Grid3D(mesh, tip_small["ids"], tsg.species_table(FIXTURE_SPECIES), 0.0, mode)
    for _ in range(20):
        g.relax(1)
    rng = np.random.default_rng(3)
    first = rng.permutation(g.n).astype(np.int32)                 # every node, shuffled
    second = np.sort(rng.choice(g.n, 257, replace=False)).astype(np.int32)
    for nodes in (first, second):
        a = g.field_o1_noalloc(nodes)
        b = g.field_o1(nodes)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), "differs on %d nodes" % len(nodes)
    # the second call's result does not depend on what the first left behind
    a = g.field_o1_noalloc(second)
    assert np.array_equal(a.view(np.uint64), g.field_o1(second).view(np.uint64))
    with pytest.raises(tsg.TsgError) as e:
        g.field_o1_noalloc(np.zeros(65537, np.int32))
    assert "staging limit" in str(e.value)
