"""GPU tests of the serving loop (tprh_server_*).

* answers: every right-hand side of every request equals the oracle's solve
  of (generated matrix, generated right-hand side) bit for bit in exact
  arithmetic, across size classes including the split path (n = 32768);
* alone vs mixed: a request's answers are identical whether it is served by
  itself or inside a stream of concurrent requests (fast arithmetic);
* admission: a budget smaller than the offered work still serves every
  request (they wait for memory), the arena never exceeds the budget, and a
  request or matrix that cannot fit is refused with TPRH_ERROR_BUDGET;
* overlap: several batches are on the device at once.
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
from tpr_amd import _lib
from tpr_amd.serve import Server

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5
SEED = 515
OKIND = {"dominant": O.GEN_DOMINANT, "weak": O.GEN_WEAK, "fixed": O.GEN_FIXED, "random_rhs": O.GEN_RANDOM_RHS}


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def bits_equal(x, y):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    same = x.view(np.uint32) == y.view(np.uint32)
    return bool(np.all(same | (np.isnan(x) & np.isnan(y))))


def oracle_answers(method, sys_id, n, nrhs, j0=0, rhs_seed=SEED, kind=O.GEN_DOMINANT):
    a, b, c, _, _ = O.gen_system(SEED, sys_id, n, kind)
    s = min(n, 256)
    return np.stack([O.solve_batch(method, a, b, c, O.gen_rhs(rhs_seed, sys_id, j0 + j, n), n, s)
                     for j in range(nrhs)])


CASES = [(16, 1), (256, 3), (1024, 8), (4096, 2), (16384, 3), (32768, 2), (512, 5), (2048, 1)]


@pytest.mark.parametrize("method", ["ptpr", "tpr"])
def test_serve_exact_matches_oracle(method):
    say("serve exact", method)
    with Server(256 << 20, method=method, arith="exact") as srv:
        reqs = []
        for k, (n, nrhs) in enumerate(CASES):
            m = srv.matrix_generate(n, seed=SEED, sys=100 + k)
            reqs.append((srv.submit(m, nrhs, rhs_seed=SEED, j0=k), m, n, nrhs, k))
        for rid, m, n, nrhs, k in reqs:
            srv.wait(rid)
            x = srv.result(rid).cpu().numpy()
            ref = oracle_answers(method, 100 + k, n, nrhs, j0=k)
            assert bits_equal(x, ref), (n, nrhs)
            srv.release(rid)
            srv.matrix_release(m)
        st = srv.stats()
        assert st["completed"] == len(CASES) and st["failed"] == 0
        assert st["bytes_in_use"] == 0


def test_serve_upload_and_caller_buffers():
    """Host-uploaded matrix, right-hand sides from a device tensor, answers
    copied into a device tensor."""
    n, nrhs = 2048, 4
    a, b, c, _, _ = O.gen_system(SEED, 7, n)
    rhs = np.stack([O.gen_rhs(9, 7, j, n) for j in range(nrhs)]).astype(np.float32)
    with Server(64 << 20, arith="exact") as srv:
        m = srv.matrix_upload(a, b, c)
        d = torch.from_numpy(rhs.ravel().copy()).cuda()
        x = torch.empty(nrhs * n, device="cuda")
        rid = srv.submit(m, nrhs, d=d, x=x)
        srv.wait(rid)
        assert srv.info(rid)["x_ptr"] is None  # answers went to x, nothing retained
        ref = np.stack([O.solve_batch("ptpr", a, b, c, rhs[j], n, 256) for j in range(nrhs)])
        assert bits_equal(x.cpu().numpy().reshape(nrhs, n), ref)
        srv.release(rid)
        srv.matrix_release(m)
        assert srv.stats()["bytes_in_use"] == 0


def _serve_all(srv, work, one_by_one):
    out = {}
    ids = []
    for k, (n, nrhs) in enumerate(work):
        m = srv.matrix_generate(n, seed=SEED, sys=1000 + k)
        rid = srv.submit(m, nrhs, rhs_seed=SEED)
        ids.append((k, rid, m))
        if one_by_one:
            srv.wait(rid)
    for k, rid, m in ids:
        srv.wait(rid)
        out[k] = srv.result(rid).cpu().numpy()
        srv.release(rid)
        srv.matrix_release(m)
    return out


def test_serve_alone_equals_mixed():
    rng = np.random.default_rng(5)
    work = [(int(1 << rng.integers(9, 16)), int(rng.integers(1, 65))) for _ in range(40)]
    say("alone vs mixed", len(work))
    with Server(1 << 30, max_inflight=4, max_batch_systems=96) as srv:
        alone = _serve_all(srv, work, True)
        mixed = _serve_all(srv, work, False)
        st = srv.stats()
    say(st)
    assert st["max_inflight_seen"] >= 2
    worst = 0.0
    for k, (n, nrhs) in enumerate(work):
        assert bits_equal(alone[k], mixed[k]), k
        if k % 8 == 0:
            ref = oracle_answers("ptpr", 1000 + k, n, min(nrhs, 4))
            worst = max(worst, max(O.rel_l2(alone[k][j], ref[j]) for j in range(min(nrhs, 4))))
    assert worst <= FAST_TOL


def test_serve_budget_admission():
    n, nrhs = 4096, 16
    req_bytes = 2 * nrhs * n * 4          # right-hand sides + answers
    budget = 6 * req_bytes                # room for a few requests at a time
    with Server(budget, max_inflight=4) as srv:
        mats = [srv.matrix_generate(n, seed=SEED, sys=2000 + k) for k in range(3)]
        xs = [torch.empty(nrhs * n, device="cuda") for _ in range(24)]
        rids = [srv.submit(mats[k % 3], nrhs, x=xs[k], rhs_seed=SEED, j0=k) for k in range(24)]
        for r in rids:
            srv.wait(r)
        st = srv.stats()
        assert st["completed"] == 24 and st["failed"] == 0
        assert st["peak_bytes"] <= budget
        for k in (0, 11, 23):
            ref = oracle_answers("ptpr", 2000 + k % 3, n, 1, j0=k)
            got = xs[k].cpu().numpy()[:n]
            assert O.rel_l2(got, ref[0]) <= FAST_TOL
        # larger than the whole budget: refused at submit
        with pytest.raises(_lib.TprError) as ei:
            srv.submit(mats[0], 200)
        assert ei.value.status == _lib.ERROR_BUDGET
        for r in rids:
            srv.release(r)
        # matrices fill the arena: the next one is refused
        big = []
        with pytest.raises(_lib.TprError) as ei:
            for k in range(1000):
                big.append(srv.matrix_generate(16384, sys=k))
        assert ei.value.status == _lib.ERROR_BUDGET
        for m in big + mats:
            srv.matrix_release(m)
        assert srv.stats()["bytes_in_use"] == 0


@pytest.mark.parametrize("kind", ["dominant", "weak", "fixed", "random_rhs"])
def test_serve_shared_generated_matrices(kind):
    """Generated matrices are filled by the first batch that uses them; later
    batches on other streams wait for that generation.  One request per
    batch, several batches in flight, each matrix released right after its
    last submit: every answer bit-identical to the oracle."""
    n, nrhs, nmat, nreq = 2048, 3, 4, 32
    say("serve shared generated", kind)
    with Server(64 << 20, arith="exact", max_inflight=4, max_batch_systems=nrhs) as srv:
        mats = [srv.matrix_generate(n, kind=kind, seed=SEED, sys=4000 + k) for k in range(nmat)]
        rids = []
        for q in range(nreq):
            rids.append(srv.submit(mats[q % nmat], nrhs, rhs_seed=SEED, j0=q))
            if q >= nreq - nmat:
                srv.matrix_release(mats[q % nmat])  # freed when its last request retires
        for q, rid in enumerate(rids):
            srv.wait(rid)
            ref = oracle_answers("ptpr", 4000 + q % nmat, n, nrhs, j0=q, kind=OKIND[kind])
            assert bits_equal(srv.result(rid).cpu().numpy(), ref), q
            srv.release(rid)
        st = srv.stats()
        assert st["batches"] == nreq and st["max_inflight_seen"] >= 2
        assert st["failed"] == 0 and st["bytes_in_use"] == 0


def test_serve_batching_window():
    """With a batching window (a batch waits for min_batch_systems queued
    right-hand sides or max_wait_us while others run) every request still
    completes, bit-identical to the oracle, in fewer batches than requests."""
    n, nrhs, nreq = 1024, 2, 48
    say("serve batching window")
    with Server(64 << 20, arith="exact", max_inflight=2, min_batch_systems=16, max_wait_us=500.0) as srv:
        m = srv.matrix_generate(n, seed=SEED, sys=4100)
        rids = [srv.submit(m, nrhs, rhs_seed=SEED, j0=nrhs * q) for q in range(nreq)]
        for q, rid in enumerate(rids):
            srv.wait(rid)
            ref = oracle_answers("ptpr", 4100, n, nrhs, j0=nrhs * q)
            assert bits_equal(srv.result(rid).cpu().numpy(), ref), q
            srv.release(rid)
        srv.matrix_release(m)
        st = srv.stats()
    assert st["completed"] == nreq and st["failed"] == 0 and st["batches"] < nreq


def test_serve_wait_any():
    """wait_any hands out every completed id exactly once, in whatever order
    the batches finish, and returns empty after its timeout when none is left."""
    say("serve wait_any")
    with Server(64 << 20, max_inflight=4, max_batch_systems=8) as srv:
        m = srv.matrix_generate(4096, seed=SEED, sys=4200)
        rids = {srv.submit(m, 4, rhs_seed=SEED, j0=4 * q) for q in range(24)}
        seen = []
        while len(seen) < len(rids):
            got = srv.poll(timeout=30.0)
            assert got, "wait_any timed out with requests outstanding"
            seen += got
        assert sorted(seen) == sorted(rids)
        assert srv.poll(timeout=0.01) == []
        for r in rids:
            srv.release(r)
        srv.matrix_release(m)


def test_serve_poll_skips_released_and_timeouts():
    """Clients that wait and release without polling: poll reports only the
    requests still held, also after more than 4096 such completions (the
    completion list is pruned of released ids); wait refuses a NaN or
    negative timeout and takes an infinite one."""
    say("serve poll after wait/release")
    with Server(64 << 20, max_inflight=4, max_batch_systems=256) as srv:
        m = srv.matrix_generate(512, seed=SEED, sys=4300)
        done = 0
        while done < 4400:
            rids = [srv.submit(m, 1, rhs_seed=SEED, j0=done + q) for q in range(200)]
            for r in rids:
                srv.wait(r)
                srv.release(r)
            done += len(rids)
        keep = [srv.submit(m, 2, rhs_seed=SEED, j0=q) for q in range(3)]
        srv.wait(keep[0], timeout=float("inf"))
        for r in keep:
            srv.wait(r)
        srv.release(keep[1])
        got = []
        while True:
            ids = srv.poll(timeout=0.05)
            if not ids:
                break
            got += ids
        assert sorted(got) == sorted([keep[0], keep[2]])
        L = _lib.lib()
        assert L.tprh_server_wait(srv._h, keep[0], float("nan")) == _lib.ERROR_INVALID_VALUE
        assert L.tprh_server_wait(srv._h, keep[0], -1.0) == _lib.ERROR_INVALID_VALUE
        srv.release(keep[0])
        srv.release(keep[2])
        srv.matrix_release(m)
        assert srv.stats()["completed"] == 4400 + 3


def test_serve_concurrent_clients():
    """Four client threads create matrices, submit, wait and release at the
    same time (the library calls release the interpreter lock): every answer
    bit-identical to the oracle, nothing left in the arena."""
    import threading
    say("serve concurrent clients")
    errors = []
    with Server(256 << 20, arith="exact", max_inflight=4, max_batch_systems=64) as srv:
        def client(c):
            try:
                rng = np.random.default_rng(900 + c)
                for q in range(8):
                    n, nrhs = int(1 << rng.integers(9, 14)), int(rng.integers(1, 5))
                    sid = 5000 + 100 * c + q
                    m = srv.matrix_generate(n, seed=SEED, sys=sid)
                    rid = srv.submit(m, nrhs, rhs_seed=SEED, j0=q)
                    srv.matrix_release(m)  # freed when the request retires
                    srv.wait(rid)
                    ref = oracle_answers("ptpr", sid, n, nrhs, j0=q)
                    if not bits_equal(srv.result(rid).cpu().numpy(), ref):
                        errors.append((c, q, n, nrhs))
                    srv.release(rid)
            except Exception as e:  # reported below, with the thread's id
                errors.append((c, repr(e)))

        threads = [threading.Thread(target=client, args=(c,)) for c in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        assert not any(t.is_alive() for t in threads)
        st = srv.stats()
    assert not errors, errors
    assert st["completed"] == 32 and st["failed"] == 0 and st["bytes_in_use"] == 0


def test_serve_destroy_drains():
    """Destroying the server with requests queued and in flight waits for all
    of them: every caller buffer then holds its answers, bit-identical to the
    oracle."""
    n, nrhs, nreq = 4096, 8, 12
    say("serve destroy drains")
    srv = Server(64 << 20, arith="exact", max_inflight=2, max_batch_systems=nrhs)
    m = srv.matrix_generate(n, seed=SEED, sys=4300)
    xs = [torch.full((nrhs * n,), float("nan"), device="cuda") for _ in range(nreq)]
    for q in range(nreq):
        srv.submit(m, nrhs, x=xs[q], rhs_seed=SEED, j0=nrhs * q)
    srv.close()
    for q in range(nreq):
        ref = oracle_answers("ptpr", 4300, n, nrhs, j0=nrhs * q)
        assert bits_equal(xs[q].cpu().numpy().reshape(nrhs, n), ref), q


def test_serve_errors():
    with Server(16 << 20) as srv:
        with pytest.raises(_lib.TprError):
            srv.matrix_generate(1000)
        m = srv.matrix_generate(1024)
        with pytest.raises(_lib.TprError):
            srv.submit(m, 0)
        with pytest.raises(_lib.TprError):
            srv.submit(m + 99, 1)
        rid = srv.submit(m, 1)
        srv.wait(rid)
        assert srv.info(rid)["state"] == "done"
        srv.release(rid)
        with pytest.raises(_lib.TprError):
            srv.release(rid)
        srv.matrix_release(m)
        with pytest.raises(_lib.TprError):
            srv.matrix_release(m)


# TPR_SERVE_SWEEP widens the sweep for a one-off run
@pytest.mark.parametrize("seed", range(int(__import__("os").environ.get("TPR_SERVE_SWEEP", "3"))))
def test_serve_random_mix_exact(seed):
    """Random request mixes under a tight budget (requests wait for memory),
    random in-flight and batch limits, matrices uploaded or generated,
    right-hand sides given or generated, answers retained or copied out:
    every answer bit-identical to the oracle."""
    rng = np.random.default_rng(500 + seed)
    method = ["ptpr", "tpr"][int(rng.integers(0, 2))]
    work = [(int(1 << rng.integers(4, 17)), int(rng.integers(1, 9))) for _ in range(int(rng.integers(10, 30)))]
    need = max(2 * n * r * 4 + (n // 256 + 1) * r * 36 + 4096 for n, r in work)
    # every matrix stays resident for the whole test (256-byte blocks); the
    # requests share what is left, which holds only a few of them at a time
    budget = int(need * rng.uniform(1.5, 4.0)) + sum(12 * n + 256 for n, _ in work)
    say("serve sweep", seed, method, len(work), budget)
    with Server(budget, method=method, arith="exact", max_inflight=int(rng.integers(1, 5)),
                max_batch_systems=int(rng.integers(8, 64))) as srv:
        # matrices first: creating one fails at once when the arena is full
        # (requests wait, matrices do not), so none is created while
        # requests hold memory
        mats = []
        for k, (n, nrhs) in enumerate(work):
            a, b, c, _, _ = O.gen_system(SEED, 3000 + k, n)
            uploaded = bool(rng.integers(0, 2))
            if uploaded:
                m = srv.matrix_upload(a, b, c)
            else:
                m = srv.matrix_generate(n, seed=SEED, sys=3000 + k)
            mats.append((m, uploaded, (a, b, c)))
        reqs = []
        for k, (n, nrhs) in enumerate(work):
            m, uploaded, (a, b, c) = mats[k]
            given = bool(rng.integers(0, 2))
            # generated right-hand sides use the matrix's generator id: 0 for an upload
            gid = 3000 + k if (given or not uploaded) else 0
            rhs = np.concatenate([O.gen_rhs(SEED + 1, gid, j, n) for j in range(nrhs)])
            d = torch.from_numpy(rhs).cuda() if given else None
            x = torch.empty(nrhs * n, device="cuda") if rng.integers(0, 2) else None
            rid = srv.submit(m, nrhs, d=d, x=x, rhs_seed=SEED + 1)
            reqs.append((k, n, nrhs, m, rid, x, (a, b, c), rhs))
        for k, n, nrhs, m, rid, x, (a, b, c), rhs in reqs:
            srv.wait(rid)
            got = (x.cpu().numpy() if x is not None else srv.result(rid).cpu().numpy()).reshape(nrhs, n)
            for j in range(nrhs):
                ref = O.solve_batch(method, a, b, c, rhs[j * n:(j + 1) * n], n, min(n, 256))
                assert bits_equal(got[j], ref), (k, n, j)
            srv.release(rid)
            srv.matrix_release(m)
        st = srv.stats()
        assert st["failed"] == 0 and st["peak_bytes"] <= budget and st["bytes_in_use"] == 0
