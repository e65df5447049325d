"""The allocation-free insert (byo_insert_batch_dev_noalloc) gives the same
container as byo_insert_batch_dev on the same batches, over two calls in a
row (the second smaller, so stale scratch contents from the first would
show), and rejects what it cannot do without allocating, container
unchanged."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import byo_amd
    assert byo_amd.device_count() > 0, "no GPU visible"
    return byo_amd


def keys_of(G):
    off, ids = G.export_csr()
    return O.graph_keys(O.Graph(off, ids))


def upd(n, count, first):
    s, d = O.update_rmat(n, count, first)
    return (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)


def test_noalloc_insert_equals_insert(B):
    g = O.rmat_graph(14)
    ref = B.CpmaGraph.from_csr(g.offsets, g.ids)
    na = B.CpmaGraph.from_csr(g.offsets, g.ids)
    E = O.graph_keys(g)
    rng = np.random.default_rng(3)

    def rnd(k):
        a = rng.integers(0, g.n, (k, 2), dtype=np.uint64)
        return (a[:, 0] << np.uint64(32)) | a[:, 1]

    # 4096 keys: new random edges, 100 edges already present and repeats;
    # then 700 new ones.  Uniform keys over RMAT-14's ~690 leaves fit their
    # leaves' slack, so neither call needs a rebalance.
    first = rnd(3400)
    first = np.concatenate([first, E[rng.integers(0, len(E), 100)], first[:596]])
    second = rnd(700)
    for batch in (first, second):
        print("noalloc insert of", len(batch), "keys", flush=True)
        d = B.DeviceBuffer(batch)
        try:
            ref.insert_batch_dev(d.ptr, len(batch))
            na.insert_batch_dev_noalloc(d.ptr, len(batch))
        finally:
            d.free()
        E = np.union1d(E, batch)
        k_ref, k_na = keys_of(ref), keys_of(na)
        assert np.array_equal(k_na, k_ref)
        assert np.array_equal(k_na, E)
        assert na.M() == ref.M() == len(E)
        assert np.array_equal(na.degrees(), ref.degrees())
    ref.close()
    na.close()


def test_noalloc_rejections_leave_container_unchanged(B):
    g = O.rmat_graph(11)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    E = O.graph_keys(g)
    big = upd(g.n, 4097, 0)
    d = B.DeviceBuffer(big)
    try:
        with pytest.raises(B.ByoError) as e:
            G.insert_batch_dev_noalloc(d.ptr, len(big))
        assert e.value.code == B.E_INVAL
    finally:
        d.free()
    bad = np.array([(np.uint64(g.n) << np.uint64(32)) | np.uint64(1)], np.uint64)
    d = B.DeviceBuffer(bad)
    try:
        with pytest.raises(B.ByoError) as e:
            G.insert_batch_dev_noalloc(d.ptr, 1)
        assert e.value.code == B.E_INVAL
    finally:
        d.free()
    # 2000 new edges out of one vertex land in one or two leaves of 512
    # slots: a rebalance is needed, which this entry point refuses
    hub = (np.uint64(9) << np.uint64(32)) | np.arange(2000, dtype=np.uint64)
    d = B.DeviceBuffer(hub)
    try:
        with pytest.raises(B.ByoError) as e:
            G.insert_batch_dev_noalloc(d.ptr, len(hub))
        assert e.value.code == B.E_STATE
    finally:
        d.free()
    assert np.array_equal(keys_of(G), E)
    assert G.M() == len(E)
    # and the same handle still takes a batch that fits
    ok = upd(g.n, 200, 77)
    d = B.DeviceBuffer(ok)
    try:
        G.insert_batch_dev_noalloc(d.ptr, len(ok))
    finally:
        d.free()
    assert np.array_equal(keys_of(G), np.union1d(E, ok))
    G.close()
