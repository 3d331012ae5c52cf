"""The exact parallel form of the triangle tree's in-place three-way partition
(TriangleAccelerator.cpp:650-701, restated in skwarka_amd/csrc/host/builders.cpp), as the device
builder computes it, checked against the sequential loop on random, degenerate and adversarial
class sequences.

The loop scans from the front; a left element is swapped to the left block's end, a right element
is swapped with the last unexamined element (which is examined next), a middle element stays. So:

* the examination sequence takes front elements in order, and after each right element takes back
  elements (from the end) up to and including the next non-right one: front element k is examined
  at step k + G(RF(k)), where RF(k) counts right elements before it and G(g) is the number of back
  elements in the first g such groups; the front elements examined are those with
  k + G(RF(k)) < n (a prefix), the rest of the array is examined from the back;
* left elements land in examination order from the front, right elements from the back;
* the middle block is a queue: a middle element appends itself, and a left element met while the
  queue is not empty moves the queue's first element to its end (swap with the block's first
  slot). With A the sequence of appends (the rho-th move appends A[rho]), the final middle block is
  the last m entries of A, each resolved to an element by following the moves back to a middle
  element (pointer jumping on the device).
"""
import numpy as np
import pytest

L, M, R = 0, 1, 2


def sequential(cls):
    """The loop itself: returns perm with perm[final position] = original index."""
    a = list(range(len(cls)))
    mid, right, i = 0, len(cls), 0
    while i < right:
        c = cls[a[i]]
        if c == L:
            a[mid], a[i] = a[i], a[mid]
            mid += 1
            i += 1
        elif c == R:
            right -= 1
            a[i], a[right] = a[right], a[i]
        else:
            i += 1
    return np.array(a), mid, right


def parallel(cls):
    """The parallel form (scans, a binary search and pointer jumping), as the device computes it."""
    cls = np.asarray(cls)
    n = len(cls)
    if n == 0:
        return np.zeros(0, np.int64), 0, 0
    is_r = cls == R
    # front prefix of right elements: RF[k] = #R in f_0..f_{k-1}
    rf = np.concatenate([[0], np.cumsum(is_r)])
    # back order b_j = element n-1-j; G[g] = back elements in the first g groups, a group ending at
    # a non-right element; groups that never end cover the rest (n)
    back_nonr = ~is_r[::-1]
    nonr_pos = np.nonzero(back_nonr)[0]
    G = np.concatenate([[0], nonr_pos + 1, np.full(n + 1, n)])
    # front elements examined: k + G(RF(k)) < n, a prefix of length K (binary search on device)
    step = np.arange(n) + G[rf[:n]]
    K = int(np.sum(step < n))
    assert np.all(step[:K] < n) and np.all(step[K:] >= n)
    E = np.empty(n, np.int64)  # examination step of each array index
    E[:K] = step[:K]
    # back element b_j (array index n-1-j, j < n-K): group g = (#non-right in b_0..b_{j-1}) + 1,
    # examined after front element k_g (the g-th right element in front order)
    nb = np.concatenate([[0], np.cumsum(back_nonr)])
    r_front_pos = np.nonzero(is_r)[0]
    for j in range(n - K):
        g = nb[j] + 1
        kg = r_front_pos[g - 1]
        E[n - 1 - j] = kg + 1 + j
    order = np.empty(n, np.int64)  # element examined at each step
    order[E] = np.arange(n)
    assert np.array_equal(np.sort(E), np.arange(n))
    ce = cls[order]
    n_l = int(np.sum(ce == L))
    n_r = int(np.sum(ce == R))
    m = n - n_l - n_r
    perm = np.empty(n, np.int64)
    l_before = np.concatenate([[0], np.cumsum(ce == L)])
    r_before = np.concatenate([[0], np.cumsum(ce == R)])
    m_before = np.concatenate([[0], np.cumsum(ce == M)])
    for q in range(n):
        if ce[q] == L:
            perm[l_before[q]] = order[q]
        elif ce[q] == R:
            perm[n - 1 - r_before[q]] = order[q]
    # the middle queue: events in examination order
    rot = (ce == L) & (m_before[:n] > 0)
    rot_before = np.concatenate([[0], np.cumsum(rot)])
    T = m + int(rot_before[n])
    src = np.full(T, -1)        # for a rotation append: the earlier append it repeats
    elem = np.full(T, -1)       # for a middle-element append: the element
    for q in range(n):
        t = m_before[q] + rot_before[q]
        if ce[q] == M:
            elem[t] = order[q]
        elif rot[q]:
            src[t] = rot_before[q]
    # pointer jumping: resolve every append to its element
    ptr = np.where(src >= 0, src, np.arange(T))
    while True:
        nxt = ptr[ptr]
        if np.array_equal(nxt, ptr):
            break
        ptr = nxt
    resolved = elem[ptr]
    r_total = int(rot_before[n])
    perm[n_l:n_l + m] = resolved[r_total:r_total + m]
    return perm, n_l, n - n_r


def check(cls):
    a, mid, right = sequential(cls)
    b, mid2, right2 = parallel(cls)
    assert (mid, right) == (mid2, right2)
    assert np.array_equal(a, b), (cls, a, b)


@pytest.mark.parametrize("seed", range(40))
def test_random_sequences(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 300))
    p = rng.dirichlet([1, 1, 1])
    check(rng.choice(3, n, p=p))


@pytest.mark.parametrize("cls", [[L] * 7, [M] * 7, [R] * 7, [R, L], [L, R], [M, L], [R, M, L], [R, R, M, L, M],
                                 [M, M, L, L, L, M, R, R, L], [R] * 3 + [M] * 3 + [L] * 3, [L, M] * 9,
                                 [M, R, L] * 11, [R, L, M] * 11, [R], [L], [M]])
def test_edge_sequences(cls):
    check(cls)


def test_exhaustive_short_sequences():
    for n in range(1, 8):
        for code in range(3 ** n):
            cls = [(code // 3 ** i) % 3 for i in range(n)]
            check(cls)
