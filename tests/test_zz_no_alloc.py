"""mte_mlp_forward_noalloc through the C ABI: the same results as
mte_mlp_forward, bit for bit, on two calls in a row (the second smaller, so
rows the first call left in the shared staging array must not leak into
it), and a clear error above the row limit."""
import numpy as np
import pytest

from mt_amd.engine import Engine, EngineError, MLP_IN, MLP_NOALLOC_ROWS, MLP_NPARAM

pytestmark = pytest.mark.gpu


def test_forward_noalloc_matches_forward():
    print("case mte_mlp_forward_noalloc vs mte_mlp_forward", flush=True)
    from mt_amd import params as P
    from tests.helpers import c1_batch
    eng = Engine(P.derive(dict(P.TEA)), c1_batch(), max_lj=64)
    rng = np.random.default_rng(8)
    eng.mlp_set_weights(rng.normal(0.0, 0.08, MLP_NPARAM).astype(np.float32))
    for n in (5000, 733):
        x = rng.normal(0.0, 0.5, (n, MLP_IN)).astype(np.float32)
        want = eng.mlp_forward(x)
        got = eng.mlp_forward_noalloc(x)
        assert np.isfinite(want).all() and np.abs(want).max() > 0
        np.testing.assert_array_equal(got, want)
    with pytest.raises(EngineError) as ei:
        eng.mlp_forward_noalloc(np.zeros((MLP_NOALLOC_ROWS + 1, MLP_IN), np.float32))
    assert ei.value.code == 1
    # a second engine on the same device shares the array and still agrees
    other = Engine(P.derive(dict(P.TEA)), c1_batch(), max_lj=64)
    other.mlp_set_weights(eng.mlp_get_weights())
    x = rng.normal(0.0, 0.5, (MLP_NOALLOC_ROWS, MLP_IN)).astype(np.float32)
    np.testing.assert_array_equal(other.mlp_forward_noalloc(x), eng.mlp_forward(x))
    other.close()
    eng.close()
