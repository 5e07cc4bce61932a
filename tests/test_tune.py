"""Autotuner host logic: the candidate grid and how a configuration is applied to an engine."""
from lagbound_amd import tune


class _Eng:
    def __init__(self):
        self.calls = []

    def set_tuning(self, *a):
        self.calls.append(("tuning", a))

    def set_option(self, name, value):
        self.calls.append(("option", name, value))


def test_candidates_and_apply():
    cs = list(tune.candidates())
    assert cs and all(c["seg_len"] >= c["heavy_edges"] for c in cs)
    assert len({tuple(sorted(c.items())) for c in cs}) == len(cs)
    e = _Eng()
    tune.apply(e, cs[0])
    assert e.calls[0] == ("tuning", (cs[0]["heavy_edges"], cs[0]["heavy_tris"], cs[0]["seg_len"], cs[0]["row_grid"]))
    assert e.calls[1] == ("option", "row_order", cs[0]["row_order"])
    small = list(tune.candidates({"heavy": [128], "seg_len": [256], "row_grid": [2048, 4096]}))
    assert len(small) == 2 * len(tune.DEFAULT_GRID["row_order"])
    assert {c["row_order"] for c in small} == {0, 1}
