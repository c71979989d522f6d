"""The flat vertex-table containers of the reference (LL_FLAT_VT: one flat
table per level; LL_ONE_VT: one flat table shared by every level;
ll_common.h:311-323, ll_mem_array.h:554-904) hold, right after each level's
freeze, exactly the vertex and edge tables of the copy-on-write container
(tests/golden/{hand_2level,rmat_s5_4level,random_48_5level}.json): the vertex
table of the new level word for word, and every older edge-table word as the
copy-on-write build has it, except that deletions by later levels have not
happened yet.  So each freeze's tables go through ll_gpu_graph_push_level
unchanged, with the same deletion patches; the reference tasks give the same
results on all three containers.  (tests/test_gpu_flat.py pushes them.)"""
import pytest

from flat_util import FLAT, LEVEL_MASK, BITS_INDEX, freezes, load, maxlev
from test_golden import load as load_cow

NEVER = LEVEL_MASK  # max-visible level of a word no level deleted


@pytest.mark.parametrize("name", FLAT)
def test_flat_tables_equal_cow_tables(name):
    doc = load(name)
    cow = load_cow(doc["source_fixture"][:-5])["reference"]
    L = cow["levels"]
    fr = doc["reference"]["freeze"]
    assert len(fr) == L == doc["reference"]["levels"]
    for k, f in enumerate(fr):
        for d in (0, 1):
            assert f["vt"][f"{d}:{k}"] == cow["vt"][f"{d}:{k}"], (name, k, d)
            for j in range(k + 1):
                got, fin = f["et"][f"{d}:{j}"], cow["et"][f"{d}:{j}"]
                assert len(got) == len(fin)
                for a, b in zip(got, fin):
                    if maxlev(b) <= k or maxlev(b) == NEVER:
                        assert a == b, (name, k, d, j)
                    else:  # deleted by a later level: not yet at freeze k
                        assert a == (b & ((1 << BITS_INDEX) - 1)) | (NEVER << BITS_INDEX)


@pytest.mark.parametrize("name", FLAT)
def test_flat_patches_are_the_deletions_of_each_level(name):
    doc = load(name)
    src = load_cow(doc["source_fixture"][:-5])
    for k, per_dir in enumerate(freezes(doc)):
        n_del = len(src["batches"][k]["del"])
        for d in (0, 1):
            _, _, ids, _ = per_dir[d]
            # each deleted edge is one word of an older level (or of none if
            # it was inserted in this level, which the fixtures avoid)
            assert ids.size <= n_del, (name, k, d)


@pytest.mark.parametrize("name", FLAT)
def test_flat_reference_tasks_equal_cow(name):
    doc = load(name)
    cow = load_cow(doc["source_fixture"][:-5])["reference"]["tasks"]
    for key, t in doc["reference"]["tasks"].items():
        assert t["values"] == cow[key]["values"], (name, key)
        assert t["return"] == cow[key]["return"], (name, key)
