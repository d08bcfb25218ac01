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
