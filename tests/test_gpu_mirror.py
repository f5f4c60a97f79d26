"""The lazily filled host mirror of an NArray (device-resident lifecycle):
element reads fetch one 64 KiB page and are then served from host memory,
element writes stay on the host until the device next uses the array (one
write-back per run of dirty pages), a device write drops the mirror, and a
pending fill answers host reads without any copy.  Values are checked against
numpy copies of the same data; the traffic against the mirror's counters."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PAGE = 1 << 16


def _delta(ph, before):
    after = ph.mirror_stats()
    return {k: after[k] - before[k] for k in after}


def test_element_reads_fetch_one_page(ph):
    a = ph.NArray.random([1000, 1000], ph.DType.F32, seed=5)
    h = a.to_numpy()
    s0 = ph.mirror_stats()
    vals = [a.get(0, j) for j in range(0, 1000, 7)]  # row 0: the first 4000 bytes
    d = _delta(ph, s0)
    assert d["pages_fetched"] == 1 and d["bytes_fetched"] == PAGE
    assert vals == [float(h[0, j]) for j in range(0, 1000, 7)]
    s1 = ph.mirror_stats()
    assert a.get(3, 999) == float(h[3, 999]) and a.get(16, 383) == float(h[16, 383])  # same page
    assert _delta(ph, s1)["pages_fetched"] == 0
    assert a.get(999, 999) == float(h[999, 999])  # the last (partial) page
    d = _delta(ph, s1)
    assert d["pages_fetched"] == 1 and d["bytes_fetched"] == 4 * 10 ** 6 - (4 * 10 ** 6 // PAGE) * PAGE


def test_element_writes_are_written_back_before_device_use(ph):
    a = ph.NArray.random([512, 512], ph.DType.F64, seed=6)  # 2 MiB: 32 pages
    h = a.to_numpy()
    s0 = ph.mirror_stats()
    for i in range(0, 512, 3):      # pages 0..31, every one dirty: one run
        a.set_element([i, i], -1.5)
        h[i, i] = -1.5
    for j in range(100):            # more writes into page 0
        a.set_element([0, j], float(j))
        h[0, j] = float(j)
    d = _delta(ph, s0)
    assert d["runs_written"] == 0 and d["pages_fetched"] == 32
    assert a.sum(acc="f64") == pytest.approx(h.sum(), rel=1e-12)  # device use: write-back first
    d = _delta(ph, s0)
    assert d["runs_written"] == 1 and d["bytes_written"] == 512 * 512 * 8
    assert np.array_equal(a.to_numpy(), h)
    # a second device use writes nothing back
    s1 = ph.mirror_stats()
    a.sum(acc="f64")
    assert _delta(ph, s1)["runs_written"] == 0
    # separate dirty runs: pages 2 and 18 (16 rows of 4096 bytes per page)
    a.set_element([32, 0], 3.0)
    h[32, 0] = 3.0
    a.set_element([288, 0], 4.0)
    h[288, 0] = 4.0
    s2 = ph.mirror_stats()
    assert np.array_equal(a.to_numpy(), h)
    assert _delta(ph, s2)["runs_written"] == 2


def test_device_writes_drop_the_mirror(ph):
    a = ph.NArray.random([300, 300], ph.DType.F64, seed=7)
    h = a.to_numpy()
    assert a.get(10, 10) == h[10, 10]
    a.map_(lambda v: v * 2.0 + 1.0)            # in place on the device
    assert a.get(10, 10) == h[10, 10] * 2.0 + 1.0
    a[["0..4", "0..4"]] = 0.25                 # device fill of a region
    assert a.get(2, 3) == 0.25 and a.get(10, 10) == h[10, 10] * 2.0 + 1.0
    b = ph.NArray.random([5, 5], ph.DType.F64, seed=8)
    a[["5..9", "5..9"]] = b                    # device scatter
    assert a.get(7, 7) == b.to_numpy()[2, 2]
    # views and mutable views read and write through the same mirror
    v = a.view("5..9", "5..9")
    assert v.get(0, 0) == b.to_numpy()[0, 0]
    mv = a.mutable_view("20..29", "20..29")
    mv.set_element([1, 1], -8.0)
    assert a.get(21, 21) == -8.0
    assert a.to_numpy()[21, 21] == -8.0


def test_pending_fill_and_host_writes(ph):
    z = ph.NArray.fill([1000, 1000], 7.0)
    s0 = ph.mirror_stats()
    assert z.get(3, 4) == 7.0 and z.get(999, 999) == 7.0
    assert _delta(ph, s0)["pages_fetched"] == 0      # answered from the fill value
    z.set_element([5, 5], 1.0)
    assert z.sum(acc="f64") == pytest.approx(7.0 * 10 ** 6 - 6.0, rel=1e-15)  # fill issued, then write-back
    h = np.full((1000, 1000), 7.0)
    h[5, 5] = 1.0
    assert np.array_equal(z.to_numpy(), h)
    # a fill with a dirty host page is not fused with a following scatter
    w = ph.NArray.fill([64, 64], 0, ph.DType.I32)
    w.set_element([0, 0], 5)
    src = ph.NArray.from_numpy(np.arange(16, dtype=np.int32).reshape(4, 4))
    w.mutable_view("8..11", "8..11")[[]] = src
    exp = np.zeros((64, 64), np.int32)
    exp[0, 0] = 5
    exp[8:12, 8:12] = np.arange(16).reshape(4, 4)
    assert np.array_equal(w.to_numpy(), exp)
