"""Pins the oracle's named temporal and ReSTIR AOV writes (oracle/aovs.py;
adapter.cpp:2598-2672) by known answers: which record field each AOV reads,
the resolved-record precedence (reuse record, else initial, else an empty
record with light 0xFFFFFFFF), the weight AOV's normalization
(weight_sum / M) / selected_target and its zero cases, and the bottom-up
row order of WriteRowsFlippedLocked."""
import numpy as np

from oracle import aovs

F32 = np.float32


def _records(h=2, w=3):
    cur = np.zeros((h, w), aovs.TEMPORAL_PIXEL)
    val = np.zeros((h, w), aovs.VALIDITY_PIXEL)
    rad = np.zeros((h, w), aovs.RADIANCE_PIXEL)
    for y in range(h):
        for x in range(w):
            cur[y, x]["motion_depth"] = (x, y, 0.5, 10 * y + x)
            val[y, x]["valid"] = (x + y) % 2
            val[y, x]["rejection"] = 3 + x
            rad[y, x]["sample_count"] = 100 + 10 * y + x
    return cur, val, rad


def test_temporal_aovs_read_their_fields_bottom_up():
    cur, val, rad = _records()
    motion = aovs.temporal_aov("lotus:motion", cur, val, rad)
    assert motion.dtype == F32 and motion.shape == (2, 3, 4)
    assert tuple(motion[0, 2]) == (2, 1, 0.5, 12)  # the top row's pixel lands in the bottom row
    assert tuple(motion[1, 0]) == (0, 0, 0.5, 0)
    assert aovs.temporal_aov("lotus:historyLength", cur, val, rad).tolist() == [[110, 111, 112], [100, 101, 102]]
    assert aovs.temporal_aov("lotus:historyLength", cur, val, None).tolist() == [[0] * 3] * 2
    assert aovs.temporal_aov("lotus:historyValidity", cur, val, rad).tolist() == [[1, 0, 1], [0, 1, 0]]
    assert aovs.temporal_aov("lotus:historyRejection", cur, val, rad).tolist() == [[3, 4, 5]] * 2


def test_restir_aovs_resolve_and_normalize():
    ini = np.zeros((1, 3), aovs.RESTIR_PIXEL)
    ini["light"] = (7, 8, 9)
    ini["weight_sum"] = (6.0, 1.0, 2.0)
    ini["selected_target"] = (0.5, 0.0, 4.0)
    ini["candidate_count"] = (3, 5, 0)
    ini["proposal"] = 0.25
    ini["contribution"] = [(1, 2, 3, 1), (0, 0, 0, 0), (4, 5, 6, 0)]
    w = aovs.restir_aov("lotus:restirWeight", ini)
    assert w[0, 0].tolist() == [6.0, 0.5, 4.0, 0.25]  # (6 / 3) / 0.5
    assert w[0, 1, 2] == 0 and w[0, 2, 2] == 0  # no target; no candidates
    assert aovs.restir_aov("lotus:restirLight", ini).tolist() == [[7, 8, 9]]
    assert aovs.restir_aov("lotus:restirCandidates", ini).tolist() == [[3, 5, 0]]
    assert aovs.restir_aov("lotus:restirSupport", ini).tolist() == [[3, 5, 0]]
    assert aovs.restir_aov("lotus:restirAge", ini).tolist() == [[0, 0, 0]]
    assert aovs.restir_aov("lotus:restirDirect", ini)[0, 2].tolist() == [4, 5, 6, 0]
    reuse = np.zeros((1, 3), aovs.RESTIR_REUSE_PIXEL)
    reuse["reservoir"] = ini
    reuse["normalization"] = (0.125, 0.5, 2.0)
    reuse["supporting_candidates"] = (11, 12, 13)
    reuse["age"] = (1, 2, 3)
    assert aovs.restir_aov("lotus:restirWeight", ini, reuse)[0, :, 2].tolist() == [0.125, 0.5, 2.0]
    assert aovs.restir_aov("lotus:restirSupport", ini, reuse).tolist() == [[11, 12, 13]]
    assert aovs.restir_aov("lotus:restirAge", ini, reuse).tolist() == [[1, 2, 3]]
    empty = aovs.restir_aov("lotus:restirLight", shape=(2, 2))
    assert empty.tolist() == [[-1, -1], [-1, -1]]
    assert aovs.restir_aov("lotus:restirWeight", shape=(2, 2)).sum() == 0
