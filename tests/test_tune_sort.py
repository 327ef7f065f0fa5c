"""The sortArray autotuner's decisions (mgpu_old_amd/tune_sort.py), on the CPU:
the built-in choice stands unless a candidate beats it by the margin, and
runs of equal choices become one table row."""
from mgpu_old_amd import tune_sort


def test_choose_keeps_builtin_within_margin():
    times = {("small", "one"): 1.00, ("large", "one"): 0.99, ("small", "xcd"): 1.2,
             ("large", "xcd"): 1.1, "builtin": 1.0}
    assert tune_sort.choose(times, ("small", "one"), 0.03) == ("small", "one")
    times[("large", "one")] = 0.90
    assert tune_sort.choose(times, ("small", "one"), 0.03) == ("large", "one")


def test_table_rows_merge_runs_and_cover_every_size():
    s, l, lx = ("small", "one"), ("large", "one"), ("large", "xcd")
    choices = {"keys": [(16, s), (17, s), (18, l), (19, l), (20, lx)],
               "kv": [(16, s), (17, l), (18, s)]}
    assert tune_sort.table_rows(choices) == [
        (17, "keys", "small", "one"), (19, "keys", "large", "one"), (31, "keys", "large", "xcd"),
        (16, "kv", "small", "one"), (17, "kv", "large", "one"), (31, "kv", "small", "one")]
