"""Host logic of the batch stage (no GPU): which autotune shapes it times."""
from __future__ import annotations

import pytest

from hulktracer_amd.workloads import _TUNE_KEYS, batch_candidates


def _row(ms, pipeline, block, coop=0):
    return dict(zip(_TUNE_KEYS, (5, block, 6, 0, 1, pipeline, 1, coop)), ms=ms)


def test_batch_candidates_keep_every_pipeline():
    # lane-refill shapes (pipeline 3) fill the table's top; the fused ones (2) must still be timed
    table = [_row(1.0 + 0.01 * i, 3, b, c) for i, (b, c) in enumerate(
        [(64, 0), (64, 16), (64, 64), (128, 0), (128, 16), (128, 64), (256, 0), (256, 64)])]
    table += [_row(2.0, 2, 64), _row(2.1, 2, 128), _row(2.2, 2, 256), _row(3.0, 1, 64)]
    cands = batch_candidates(table, top=6)
    pipes = [k[_TUNE_KEYS.index("pipeline")] for k in cands]
    assert len(cands) == 6 and len(set(cands)) == 6
    assert pipes.count(3) == 3 and pipes.count(2) == 3
    assert cands[0] == tuple(int(table[0][f]) for f in _TUNE_KEYS)  # the fastest shape first


def test_batch_candidates_fill_from_one_pipeline_and_prefer():
    table = [_row(1.0 + 0.01 * i, 3, 64 * (1 + i % 4), 16 * (i // 4)) for i in range(8)]
    pref = [(5, 512, 6, 0, 1, 3, 1, 64)]
    cands = batch_candidates(table, top=6, prefer=pref)
    assert len(cands) == 7 and cands[-1] == pref[0]  # six from the one pipeline, then the preferred shape
    with pytest.raises(ValueError):
        batch_candidates([], top=6)
