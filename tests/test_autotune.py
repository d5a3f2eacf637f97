"""The autotuner's search space (no GPU needed): settings are complete and the two stages
cover the binning shapes, then refine the fastest one."""
import pytest

pytest.importorskip("torch")


def test_autotune_staged_search():
    """The default search: 20 binning shapes (with and without the vertex pass), then 48
    refinements of the fastest one."""
    from softwarerenderer_amd import autotune as T
    first = list(T.staged_configs())
    assert len(first) == 20 and all(set(c) == set(T.KNOBS) for c in first)
    shape = first[7]
    second = list(T.staged_configs(shape))
    assert len(second) == 48
    assert all(c["bin_size"] == shape["bin_size"] and c["vertex_pass"] == shape["vertex_pass"]
               for c in second)
