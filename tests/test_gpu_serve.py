"""Serving loop (r3dg_server_start / r3dg_submit / r3dg_wait): pairs that
arrive while others run share the loop's steps (continuous batching), and
each result is bit-identical to the pair matched alone (r3dg_match), for
uploaded images and for cached views reused across partners."""
import threading
import time

import numpy as np
import pytest

from regard3d_mvs_amd import StereoOptions, synth

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    bad = a.view(np.uint32) != b.view(np.uint32)
    assert int(bad.sum()) == 0, f"{int(bad.sum())} of {a.size} differ"


def views_and_requests():
    """Three views, each the left of one pair and the right of another (the
    partner image of a pair is the right image of a generated pair)."""
    gens = [synth.make_pair(w, h, span, 900 + i)
            for i, (w, h, span) in enumerate([(300, 180, 48), (256, 200, 64), (180, 120, 32)])]
    reqs = []
    for i, p in enumerate(gens):
        reqs.append(dict(left=p.left, right=p.right, dmin=p.dmin, nd=p.nd))
    # A second partner for view 0 (its own right image, cropped narrower).
    reqs.append(dict(left=gens[0].left, right=gens[0].right[:, :260].copy(), dmin=0, nd=48))
    return gens, reqs


@pytest.mark.parametrize("budget", [0, 4 << 20])
def test_serve_alone_equals_in_mix(budget):
    """With the default budget all pairs overlap; with a 4 MiB budget only
    one pair is admitted at a time."""
    from regard3d_mvs_amd.native import Matcher
    gens, reqs = views_and_requests()
    o = StereoOptions()
    with Matcher(0) as m:
        alone = [m.match(q["left"], q["right"], None, None, q["dmin"], q["nd"], o) for q in reqs]
        m.serve_start(o, hbm_budget=budget)
        tickets = []
        for k, q in enumerate(reqs):
            tickets.append(m.submit(q["left"].shape, q["dmin"], q["nd"], left=q["left"], right=q["right"]))
            if k == 1:
                time.sleep(0.01)  # later arrivals join a loop that is running
        got = [m.wait(t) for t in tickets]
        m.serve_stop()
    for g, a in zip(got, alone):
        assert g.cells == a.cells and g.levels == a.levels
        same_bits(g.disparity, a.disparity)
        same_bits(g.confidence, a.confidence)
        assert g.seconds > 0


def test_serve_cached_views_and_threads():
    """Views put once and named by several requests submitted from two
    threads give the same maps as uploads matched alone."""
    from regard3d_mvs_amd.native import Matcher
    gens, reqs = views_and_requests()
    o = StereoOptions()
    with Matcher(0) as m:
        for i, p in enumerate(gens):
            m.view_put(10 + i, p.left)
            m.view_put(20 + i, p.right)
        alone = [m.match(p.left, p.right, None, None, p.dmin, p.nd, o) for p in gens]
        m.serve_start(o)
        tickets = [None] * 6

        def client(k0):
            for k in range(k0, 6, 2):
                i = k % 3
                tickets[k] = m.submit(gens[i].left.shape, gens[i].dmin, gens[i].nd, left_view=10 + i,
                                      right_view=20 + i)

        th = [threading.Thread(target=client, args=(k,)) for k in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        got = [m.wait(t) for t in tickets]
        m.view_evict(10)
        m.serve_stop()
    for k, g in enumerate(got):
        same_bits(g.disparity, alone[k % 3].disparity)
        same_bits(g.confidence, alone[k % 3].confidence)


def test_serve_views_put_while_running():
    """Views put while the loop runs go up through pinned staging between its
    steps: uint8 views widened on the device, uint16 views as they are, a view
    with a mask.  Each served pair equals the same images matched alone."""
    from regard3d_mvs_amd.native import Matcher
    p = synth.make_pair(1280, 900, 64, 931)   # uint16 views of 2.3 MB
    left8 = (p.left >> 8).astype(np.uint8)   # uint8 views of 1.2 MB
    right8 = (p.right >> 8).astype(np.uint8)
    mask = np.full(p.left.shape, 255, np.uint8)
    mask[100:300, 200:500] = 0
    o = StereoOptions()
    with Matcher(0) as m:
        alone = [m.match(p.left, p.right, None, None, p.dmin, p.nd, o),
                 m.match(left8, right8, None, None, p.dmin, p.nd, o),
                 m.match(left8, right8, mask, None, p.dmin, p.nd, o)]
        m.serve_start(o)
        try:
            m.view_put(1, p.left)
            m.view_put(2, p.right)
            m.view_put(3, left8)
            m.view_put(4, right8)
            m.view_put(5, left8, mask)
            tickets = [m.submit(p.left.shape, p.dmin, p.nd, left_view=1, right_view=2),
                       m.submit(p.left.shape, p.dmin, p.nd, left_view=3, right_view=4),
                       m.submit(p.left.shape, p.dmin, p.nd, left_view=5, right_view=4)]
            got = [m.wait(t) for t in tickets]
        finally:
            m.serve_stop()
    for g, a in zip(got, alone):
        assert g.cells == a.cells
        same_bits(g.disparity, a.disparity)
        same_bits(g.confidence, a.confidence)


def test_serve_admission_counts_pairs_from_views():
    """Pairs whose images all come from cached views are admitted by their
    estimated footprint too: with an admission budget of one byte the loop
    runs them strictly one at a time, and every map equals the pair alone."""
    from regard3d_mvs_amd.native import Matcher
    gens, _ = views_and_requests()
    o = StereoOptions()
    with Matcher(0) as m:
        for i, p in enumerate(gens):
            m.view_put(10 + i, p.left)
            m.view_put(20 + i, p.right)
        alone = [m.match(p.left, p.right, None, None, p.dmin, p.nd, o) for p in gens]
        m.serve_start(o, hbm_budget=1)
        try:
            tickets = [m.submit(gens[k % 3].left.shape, gens[k % 3].dmin, gens[k % 3].nd, left_view=10 + k % 3,
                                right_view=20 + k % 3) for k in range(6)]
            most = 0
            while True:
                st = m.server_stats()
                most = max(most, st["active"])
                if st["completed"] == len(tickets):
                    break
                time.sleep(0.0005)
            got = [m.wait(t) for t in tickets]
        finally:
            m.serve_stop()
    assert most <= 1, most
    for k, g in enumerate(got):
        same_bits(g.disparity, alone[k % 3].disparity)


def test_serve_errors():
    from regard3d_mvs_amd.native import Matcher, R3dgError
    gens, reqs = views_and_requests()
    with Matcher(0) as m:
        with pytest.raises(R3dgError):
            m.submit((10, 10), 0, 16, left_view=1, right_view=2)  # no loop running
        m.serve_start(StereoOptions())
        t = m.submit(gens[0].left.shape, 0, 32, left_view=77, right_view=78)  # unknown views
        with pytest.raises(R3dgError):
            m.wait(t)
        t2 = m.submit(reqs[2]["left"].shape, reqs[2]["dmin"], reqs[2]["nd"], left=reqs[2]["left"],
                      right=reqs[2]["right"])
        assert m.wait(t2) is not None
        m.serve_stop()


def test_serve_under_memory_pressure():
    """A context budget far below what the admitted pairs would need with
    their volumes whole: the steps chunk the volumes to the budget and every
    pair completes, bit-identical to alone."""
    from regard3d_mvs_amd.native import Matcher
    pairs = [synth.make_pair(480, 320, 64, 950 + i) for i in range(6)]
    o = StereoOptions()
    with Matcher(0, hbm_budget=160 << 20) as m:
        alone = [m.match(p.left, p.right, None, None, p.dmin, p.nd, o) for p in pairs]
        m.serve_start(o, hbm_budget=64 << 30)  # admission alone would take all six at once
        try:
            tickets = [m.submit(p.left.shape, p.dmin, p.nd, left=p.left, right=p.right) for p in pairs]
            got = [m.wait(t) for t in tickets]
            st = m.server_stats()
        finally:
            m.serve_stop()
    for g, a in zip(got, alone):
        same_bits(g.disparity, a.disparity)
        same_bits(g.confidence, a.confidence)
    assert st["submitted"] == st["completed"] == len(pairs) and st["failed"] == 0, st
    assert st["pending"] == 0 and st["active"] == 0 and st["steps"] >= 2, st


@pytest.mark.parametrize("after", [0, 40])
def test_serve_requeues_pairs_when_a_step_runs_out(after):
    """A device allocation that fails inside a step of several pairs: the pairs
    go back to the queue, are admitted again and complete bit-identical to
    alone; none fails."""
    from regard3d_mvs_amd.native import Matcher
    pairs = [synth.make_pair(480, 320, 64, 970 + i) for i in range(6)]
    o = StereoOptions()
    with Matcher(0) as m:
        alone = [m.match(p.left, p.right, None, None, p.dmin, p.nd, o) for p in pairs]
        m.serve_start(o)
        try:
            m.inject_step_failure(after)
            tickets = [m.submit(p.left.shape, p.dmin, p.nd, left=p.left, right=p.right) for p in pairs]
            got = [m.wait(t) for t in tickets]
            st = m.server_stats()
        finally:
            m.serve_stop()
    for g, a in zip(got, alone):
        same_bits(g.disparity, a.disparity)
        same_bits(g.confidence, a.confidence)
    assert st["completed"] == len(pairs) and st["failed"] == 0, st
    assert st["requeued"] >= 2 and st["admission_scale"] == 0.75, st
