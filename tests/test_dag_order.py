"""CPU checks of the task-graph Cholesky's claim order, as the kernel's workers decode it
(rgf_debug_dag_order runs the same host/device decoding): every task of the blocked right-
looking Cholesky (and of L^-1) appears exactly once, and the order is one that a single worker
beside the chain can always execute -- the workers wait on the tasks they claim, so a claim
order that lists a task before one it depends on could stall the launch until its bounded
wait expires.  The dependencies are the waits of potrf_dag_kernel: the chain step k needs
tile (k, k) with panels 0..k-2 applied, then L_{k+1,k-1} and tile (k+1, k) with panels
0..k-2; TRSM(i, g) needs X_g (the chain), L_{i,g-1} and tile (i, g) with panels 0..g-2;
UPDATE(i, j, p) needs L_ip, L_jp and tile (i, j) with panels 0..p-1; XACC(i, j, k) needs
L_ik, X_kj (the chain's X_jj for k = j) and Y_ij with terms j..k-1."""
import ctypes as C

import numpy as np
import pytest

from rgf_amd import _lib


def order(T, with_x, xlag, promo):
    lib = _lib.load()
    n = lib.rgf_debug_dag_order(T, int(with_x), xlag, int(promo), None, 0)
    assert n >= 0
    out = (C.c_int * max(5 * n, 1))()
    assert lib.rgf_debug_dag_order(T, int(with_x), xlag, int(promo), out, n) == n
    return [tuple(out[5 * t: 5 * t + 5]) for t in range(n)]


def expected_tasks(T, with_x):
    """The task set by definition: TRSM(i, g) for i >= g + 2 (i = g + 1 is the chain's),
    UPDATE(i, j, p) for i >= j >= p + 2, XACC(i, j, k) for i > k >= j."""
    s = set()
    for g in range(T):
        for i in range(g + 2, T):
            s.add((1, i, g, g))
    for p in range(T):
        for j in range(p + 2, T):
            for i in range(j, T):
                s.add((2, i, j, p))
    if with_x:
        for k in range(T - 1):
            for j in range(k + 1):
                for i in range(k + 1, T):
                    s.add((6, i, j, k))
    return s


def runs_to_completion(T, tasks, workers):
    """The chain and `workers` workers; a worker claims the next task and waits for its inputs."""
    xdone, ldone, cnt, acnt, xrdy = set(), set(), {}, {}, set()
    chain_k, phase, nxt = 0, 0, 0
    held = [None] * workers

    def ready(tk):
        kind, i, j, p = tk[:4]
        if kind == 1:
            g = j
            return g in xdone and (g < 1 or (i, g - 1) in ldone) and (g < 2 or cnt.get((i, g), 0) >= g - 1)
        if kind == 2:
            return (i, p) in ldone and (j == i or (j, p) in ldone) and (p < 1 or cnt.get((i, j), 0) >= p)
        k = p
        return ((i, k) in ldone and ((j in xdone) if k == j else ((k, j) in xrdy))
                and (k <= j or acnt.get((i, j), 0) >= k - j))

    def run(tk):
        kind, i, j, p = tk[:4]
        if kind == 1:
            ldone.add((i, j))
        elif kind == 2:
            assert cnt.get((i, j), 0) == p   # panels reach each tile in increasing order
            cnt[(i, j)] = p + 1
        elif p < i - 1:
            assert acnt.get((i, j), 0) == p - j   # terms reach each Y_ij in increasing k
            acnt[(i, j)] = p - j + 1
        else:
            xrdy.add((i, j))

    while True:
        moved = False
        if chain_k < T:
            k = chain_k
            if phase == 0 and (k < 2 or cnt.get((k, k), 0) >= k - 1):
                xdone.add(k)
                if k >= 1:
                    ldone.add((k, k - 1))
                phase, moved = 1, True
            if phase == 1 and (k + 1 >= T or ((k < 1 or (k + 1, k - 1) in ldone)
                                              and (k < 2 or cnt.get((k + 1, k), 0) >= k - 1))):
                chain_k, phase, moved = chain_k + 1, 0, True
        for w in range(workers):
            if held[w] is None and nxt < len(tasks):
                held[w], nxt, moved = tasks[nxt], nxt + 1, True
            if held[w] is not None and ready(held[w]):
                run(held[w])
                held[w], moved = None, True
        if not moved:
            return chain_k == T and nxt == len(tasks) and all(h is None for h in held)


@pytest.mark.parametrize("with_x", [False, True])
@pytest.mark.parametrize("xlag,promo", [(0, False), (2, True), (0, True), (5, False), (16, True)])
def test_claim_order_is_complete_and_executable(with_x, xlag, promo):
    for T in list(range(1, 13)) + [23, 45]:
        tasks = order(T, with_x, xlag, promo)
        keys = [t[:4] for t in tasks]
        assert len(keys) == len(set(keys)), T
        assert set(keys) == expected_tasks(T, with_x), T
        # generations never decrease along the order
        assert all(a[4] <= b[4] for a, b in zip(tasks, tasks[1:])), T
        for workers in (1, 3):
            assert runs_to_completion(T, tasks, workers), (T, workers)


def test_claim_order_moves_the_chain_inputs_early():
    """With promo the chain's two updates of step g + 2 are claimed in generation g, before
    any task of generation g + 1; without it they are generation g + 1's."""
    T = 10
    for promo, gen_of in ((True, lambda g: g), (False, lambda g: g + 1)):
        tasks = order(T, True, 2, promo)
        where = {t[:4]: t[4] for t in tasks}
        for g in range(T - 3):
            assert where[(2, g + 2, g + 2, g)] == gen_of(g)
            assert where[(2, g + 3, g + 2, g)] == gen_of(g)


def test_bad_arguments():
    lib = _lib.load()
    assert lib.rgf_debug_dag_order(0, 1, 0, 0, None, 0) == -1
    assert lib.rgf_debug_dag_order(5, 1, 17, 0, None, 0) == -1
    assert lib.rgf_debug_dag_order(5, 1, 0, 0, None, 3) == -1
