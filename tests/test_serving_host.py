"""Host-side bookkeeping of the serving loop: the arena range allocator."""
import numpy as np

from thejoker_amd.serving import RangeAllocator


def _check(alloc, held):
    """The free list and the held ranges tile [0, capacity) exactly."""
    cover = np.zeros(alloc.capacity, dtype=np.int64)
    for a, b in alloc.free:
        assert a < b
        cover[a:b] += 1
    for a, n in held.values():
        cover[a:a + n] += 1
    assert (cover == 1).all()
    assert all(alloc.free[k][1] < alloc.free[k + 1][0] for k in range(len(alloc.free) - 1))


def test_alloc_release_merge():
    al = RangeAllocator(100)
    a = al.alloc(30)
    b = al.alloc(30)
    c = al.alloc(40)
    assert (a, b, c) == (0, 30, 60) and al.alloc(1) is None
    al.release(b, 30)
    assert al.free == [(30, 60)]
    al.release(a, 30)
    assert al.free == [(0, 60)]
    al.release(c, 40)
    assert al.free == [(0, 100)] and al.used() == 0


def test_extend_in_place_and_take():
    al = RangeAllocator(100)
    a = al.alloc(10)
    b = al.alloc(10)
    assert not al.extend(a, 10, 20)        # b sits right after a
    assert al.extend(b, 10, 50)
    assert al.free == [(60, 100)]
    assert not al.extend(b, 50, 91)
    al.release(a, 10)
    al.take(2, 5)
    assert al.free == [(0, 2), (7, 10), (60, 100)]
    try:
        al.take(5, 5)
    except ValueError:
        pass
    else:
        raise AssertionError("took a range that is not free")


def test_random_operations_keep_the_tiling():
    rng = np.random.default_rng(0)
    al = RangeAllocator(1000)
    held = {}
    for step in range(3000):
        op = rng.integers(4)
        if op == 0 or not held:
            n = int(rng.integers(1, 80))
            s = al.alloc(n)
            if s is not None:
                held[step] = (s, n)
        elif op == 1:
            k = list(held)[rng.integers(len(held))]
            al.release(*held.pop(k))
        elif op == 2:
            k = list(held)[rng.integers(len(held))]
            s, n = held[k]
            m = n + int(rng.integers(1, 50))
            if al.extend(s, n, m):
                held[k] = (s, m)
        else:
            # move: release, allocate elsewhere, or take the old range back
            k = list(held)[rng.integers(len(held))]
            s, n = held[k]
            m = n + int(rng.integers(1, 50))
            al.release(s, n)
            t = al.alloc(m)
            if t is None:
                al.take(s, n)
            else:
                held[k] = (t, m)
        _check(al, held)


def test_share_budget():
    from thejoker_amd.serving import share_budget
    # no budget pressure: everyone gets its whole remainder
    assert list(share_budget([10, 20, 30], [1, 2, 3], 10_000, 1)) == [10, 20, 30]
    # small requests finish, the large ones split what is left equally (in work)
    out = share_budget([5, 1000, 1000], [10, 10, 20], 6050, 1)
    assert out[0] == 5
    assert out[1] * 10 == 3000 and out[2] * 20 == 3000
    # the minimum chunk wins over a tiny share
    assert list(share_budget([1000, 1000], [100, 100], 10, 64)) == [64, 64]
    # random cases: never more than a remainder, at least min(rem, min_rows),
    # and within budget whenever the minimum chunk does not force more
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = int(rng.integers(1, 20))
        rem = rng.integers(1, 10**6, n)
        M = rng.integers(3, 200, n)
        budget = int(rng.integers(10**4, 10**9))
        out = share_budget(rem, M, budget, 256)
        assert (out <= rem).all() and (out >= np.minimum(rem, 256)).all()
        forced = np.minimum(rem, 256) * M
        assert (out * M).sum() <= max(budget, 0) + forced.sum()


def test_share_budget_srpt():
    """Shortest remaining work first: the smallest rounds complete, the
    budget left goes whole to the request nearest completion, the others
    take the minimum chunk; costs per row may be fractional (a probed
    round's row costs less than its epochs)."""
    from thejoker_amd.serving import share_budget
    out = share_budget([5, 1000, 1000], [10, 10, 20], 6050, 1, policy="srpt")
    assert out[0] == 5 and out[1] == 600 and out[2] == 1
    out = share_budget([5, 1000, 1000], [10, 10, 20], 6050, 64, policy="srpt")
    assert out[0] == 5 and out[1] == 600 and out[2] == 64
    out = share_budget([100, 1000], [2.5, 1.0], 10_000, 1, policy="srpt")
    assert list(out) == [100, 1000]
    rng = np.random.default_rng(2)
    for _ in range(200):
        n = int(rng.integers(1, 20))
        rem = rng.integers(1, 10**6, n)
        M = rng.uniform(3.0, 200.0, n)
        budget = int(rng.integers(10**4, 10**9))
        out = share_budget(rem, M, budget, 256, policy="srpt")
        assert (out <= rem).all() and (out >= np.minimum(rem, 256)).all()
        forced = np.minimum(rem, 256) * M
        assert (out * M).sum() <= max(budget, 0) + forced.sum() + 1e-6
        # past the first request (in order of remaining work) left incomplete,
        # the others take only the minimum chunk (what that one leaves, under
        # one of its rows' cost, buys no later request 256 rows at M >= 3)
        order = np.argsort(rem * M, kind="stable")
        short = np.flatnonzero(out[order] < rem[order])
        if len(short):
            rest = order[short[0] + 1:]
            assert (out[rest] == np.minimum(rem[rest], 256)).all()


def zero_uniform_rng(k, inc_seed=5):
    """A numpy Generator (PCG64) whose k-th uniform (0-based) is exactly 0:
    the state after k + 1 LCG steps is 0, whose XSL-RR output is 0."""
    inc = np.random.default_rng(inc_seed).bit_generator.state["state"]["inc"]
    mult = 0x2360ED051FC65DA44385DF649FCCF645
    inv = pow(mult, -1, 1 << 128)
    s = 0
    for _ in range(k + 1):
        s = ((s - inc) * inv) % (1 << 128)
    # a fixed seed sequence: spawned child generators must be reproducible too
    bg = np.random.PCG64(inc_seed)
    bg.state = {"bit_generator": "PCG64", "state": {"state": s, "inc": inc},
                "has_uint32": 0, "uinteger": 0}
    return np.random.Generator(bg)


def test_zero_uniform_rng_construction():
    for k in (0, 1, 37, 1000):
        u = zero_uniform_rng(k).uniform(size=k + 2)
        assert u[k] == 0.0 and (np.delete(u, k) > 0).all()


def test_margins_cover_the_uniforms():
    """The settling margins: exp(-40) is below the smallest positive uniform
    2^-53, and exp underflows to exactly 0 below -745.14 (so -746 and -750)."""
    from thejoker_amd.batch import LIVE_MARGIN, LIVE_MARGIN_NONZERO_U
    assert np.exp(-LIVE_MARGIN_NONZERO_U) < 2.0 ** -53
    assert np.exp(-746.0) == 0.0 and np.exp(-LIVE_MARGIN) == 0.0
    assert np.exp(-745.0) > 0.0
