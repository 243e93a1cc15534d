"""The shipped tuning table (fastwarp_amd/tuning/gfx950.json, written by `python -m
fastwarp_amd.autotune`): one flat JSON object whose entries are "<knob>@<size class>" with integer
values among the autotuner's candidates; the autotuner's size classes are the library's
(floor(log2(size)); the library's own parser and classes are pinned in tests/cpp/host_checks.cpp)."""
from __future__ import annotations

import json
from pathlib import Path

from fastwarp_amd.autotune import CANDIDATES, OUT, size_class


def test_size_classes():
    assert [size_class(n) for n in (1, 2, 3, 4, 1023, 1024)] == [0, 1, 1, 2, 9, 10]
    assert size_class(182 * 218 * 182) == 22 and size_class(96 * 112 * 96) == 19


def test_shipped_table_holds_only_candidate_values():
    assert OUT == Path(__file__).resolve().parent.parent / "fastwarp_amd" / "tuning" / "gfx950.json"
    table = json.loads(OUT.read_text())
    allowed = {}
    for cands in CANDIDATES.values():
        for cand in cands:
            for q in range(0, len(cand), 2):
                allowed.setdefault(cand[q], set()).add(cand[q + 1])
    entries = {k: v for k, v in table.items() if not k.startswith("_")}
    assert entries and table["_command"] == "python -m fastwarp_amd.autotune"
    for key, value in entries.items():
        knob, cls = key.split("@")
        assert knob in allowed and value in allowed[knob], key
        assert 0 <= int(cls) < 40 and isinstance(value, int), key
