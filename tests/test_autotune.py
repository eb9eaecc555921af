"""The autotuner's knobs and persisted table (host side only: the setters
touch no GPU)."""
import json

import pytest

from zccl_amd import autotune


@pytest.fixture
def defaults():
    yield
    for op in autotune.OPS:
        autotune.reset(op)


def test_keys_classify_operands_and_sizes():
    assert [autotune.k_range(k) for k in (1, 2, 3, 4, 5, 8, 9, 16, 17, 32, 40)] == \
        ["1", "2", "3-4", "3-4", "5-8", "5-8", "9-16", "9-16", "17-32", "17-32", "33+"]
    assert autotune.size_class(1 << 24) == "2^24"
    assert autotune.size_class((1 << 24) + 5) == "2^24"
    assert autotune.key("homsum", 36, 8, 1 << 24) == "homsum/L36/k5-8/n2^24"


def test_knobs_set_read_back_and_reset(defaults):
    autotune.set_knobs("homsum", {"layout": 1, "threads": 128, "stage": 32768, "two_pass": 4096})
    assert autotune.get_knobs("homsum") == {"threads": 128, "layout": 1, "stage": 32768, "two_pass": 4096}
    autotune.set_knobs("decompress", {"layout": 1, "threads": 256, "groups": 3})
    assert autotune.get_knobs("decompress") == {"threads": 256, "layout": 1, "groups": 3}
    autotune.reset("homsum")
    autotune.reset("decompress")
    assert autotune.get_knobs("homsum")["layout"] == 0 and autotune.get_knobs("homsum")["two_pass"] == 1024
    assert autotune.get_knobs("decompress") == {"threads": 512, "layout": 0, "groups": 4}


@pytest.mark.parametrize("op,cfg", [("decompress", {"groups": 6}), ("homsum", {"stage": 100}),
                                    ("compress", {"threads": 96}), ("compress", {"layout": 2})])
def test_knobs_reject_out_of_range(defaults, op, cfg):
    with pytest.raises(Exception):
        autotune.set_knobs(op, cfg)


def test_tune_keeps_fastest_and_table_round_trip(defaults, tmp_path):
    # a run whose time depends on the knobs: the tuner must pick the minimum
    cost = {"layout=0,two_pass=1024": 3.0, "layout=0,two_pass=1099511627776": 2.0}

    def run():
        k = autotune.get_knobs("homsum")
        if k["layout"] == 1:
            return 5.0 + k["threads"] / 1000 + k["stage"] / 1e6
        return cost[autotune.label({"layout": 0, "two_pass": k["two_pass"]})]

    res = autotune.tune("homsum", run)
    assert res["config"] == {"layout": 0, "two_pass": 1 << 40}
    assert autotune.get_knobs("homsum")["two_pass"] == 1 << 40
    assert len(res["median_s"]) == len(autotune.SPACES["homsum"])
    table = {}
    autotune.record(table, "homsum", 36, 8, 1 << 24, res, "test")
    path = autotune.save(table, str(tmp_path / "t.json"))
    back = autotune.load(path)
    assert json.loads(open(path).read()) == back
    autotune.reset("homsum")
    assert autotune.apply(back, "homsum", 36, 7, (1 << 24) - 3) == {"layout": 0, "two_pass": 1 << 40}
    assert autotune.get_knobs("homsum")["two_pass"] == 1 << 40
    assert autotune.apply(back, "homsum", 36, 1, 1 << 24) is None


def test_committed_table_names_known_knobs():
    table = autotune.load(autotune.DEFAULT_TABLE)
    for k, e in table.items():
        op = k.split("/")[0]
        assert op in autotune.OPS
        assert set(e["config"]) <= {"threads", "layout", "stage", "two_pass", "groups"}
