"""The allocation-free entry point plis_mlp_forward_noalloc against
plis_mlp_forward, through the C ABI: two calls in a row on the same
context (the second with fewer rows, so stale scratch contents from the
first would show), both activations, and the size limit."""
import numpy as np
import pytest

from pliscuda_amd import capi, learned
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("act", [1, 0])
def test_noalloc_forward_equals_forward(act):
    net, topo, x0 = H.small_bruss()
    ctx = capi.Context(net, [topo], [x0], seed=3)
    rng = np.random.default_rng(12)
    norm = np.concatenate([np.ones(12), 2 * np.ones(12), [0.0, 1.0]]).astype(np.float32)
    w = learned.init_weights(4)
    ctx.mlp_set(w, norm)
    ctx.mlp_activation(act)
    for n in (16384, 1000):
        f = (rng.normal(size=(n, 12)) * 2 + 1).astype(np.float32)
        ref = ctx.mlp_forward(f)
        got = ctx.mlp_forward_noalloc(f)
        assert got.shape == (n, 2)
        assert np.array_equal(got, ref), (n, np.abs(got - ref).max())
    with pytest.raises(RuntimeError, match="at most 16384 rows"):
        ctx.mlp_forward_noalloc(np.zeros((16385, 12), np.float32))
    ctx.close()
