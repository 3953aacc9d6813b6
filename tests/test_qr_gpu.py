"""GPU parity of Householder QR, Q application, triangular solve and least
squares against the FP64 oracle and the reference suite's golden values
(test/testTensor.cu X:1012-1039, 1433-1535).

Tolerances: FP64 kernels agree with the oracle to 1e-11 (relative to the
matrix scale); FP32 to the reference suite's PRECISION_LOW (1e-4)."""
import numpy as np
import pytest
import torch

import golden_values as G
from gputils_amd import api

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).cuda()


@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_least_squares_batched_golden(dtype, eps):
    """X:1012-1039: square 2x2x3 systems, ||A x - b|| < eps."""
    A = G.cm(G.LS1_A, 2, 2, 3)
    b = np.array(G.LS1_B).reshape(3, 1, 2)
    tA, tb = dev(api.from_matrices(torch.from_numpy(A)).numpy(), dtype), dev(b, dtype)
    api.gels_batched(tA, tb)
    x = tb.cpu().double().numpy()
    for k in range(3):
        assert np.linalg.norm(A[k] @ x[k, 0] - b[k, 0]) < eps * 10


@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_qr_least_squares_golden_residual(dtype, eps):
    """X:1496-1535 via geqrf + ormqr(Q^T) + trsm, as QRFactoriser does."""
    A = np.array(G.QRLS_A_ROWMAJOR).reshape(1, 4, 3)
    b = np.array(G.QRLS_B).reshape(1, 1, 4)
    tA = dev(api.from_matrices(torch.from_numpy(A)).numpy(), dtype)
    tau = api.geqrf_batched(tA)
    tb = dev(b, dtype)
    api.ormqr_batched(tA, tau, tb, trans=True)
    api.trsm_upper_batched(tA, tb)
    x = tb.cpu().double().numpy()[0, 0, :3]
    rel = 1e-6 if dtype == torch.float32 else 1e-12
    assert abs(np.linalg.norm(A[0] @ x - b[0, 0]) - G.QRLS_RESIDUAL) < max(eps, rel * G.QRLS_RESIDUAL)


@pytest.mark.parametrize("m,n", [(4, 3), (20, 3), (7, 7), (64, 64), (200, 48), (600, 30)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_qr_reconstructs_and_matches_oracle(oracle, m, n, dtype):
    """X:1433-1490 (QR - A ~ 0) plus element-wise agreement of R, v and tau
    with the oracle's dgeqr2 (FP64)."""
    k = 3
    r = np.random.default_rng(m * 31 + n)
    A = r.uniform(-100, 100, (k, m, n))
    if dtype == torch.float32:
        A = A.astype(np.float32).astype(np.float64)
    tA = dev(api.from_matrices(torch.from_numpy(A)).numpy(), dtype)
    tau = api.geqrf_batched(tA)
    QR = api.to_matrices(tA).cpu().double().numpy()
    taus = tau.cpu().double().numpy()
    # Q = H_0..H_{n-1} applied to the identity (m x n), as getQR does
    I = np.zeros((k, n, m))
    for b in range(k):
        I[b, :, :n] = np.eye(n)
    tQ = dev(I, dtype)
    api.ormqr_batched(tA, tau, tQ, trans=False)
    Q = api.to_matrices(tQ).cpu().double().numpy()
    tol = 1e-4 if dtype == torch.float32 else 1e-11
    for b in range(k):
        R = np.triu(QR[b][:n, :n])
        assert np.linalg.norm(Q[b] @ R - A[b]) <= tol * np.linalg.norm(A[b])
        assert np.linalg.norm(Q[b].T @ Q[b] - np.eye(n)) <= tol * n
        if dtype == torch.float64:
            oqr, otau = oracle.geqrf(A[b])
            scale = np.abs(oqr).max()
            assert np.max(np.abs(np.triu(QR[b][:n]) - np.triu(oqr[:n]))) <= 1e-11 * scale
            assert np.max(np.abs(np.tril(QR[b], -1) - np.tril(oqr, -1))) <= 1e-10
            assert np.max(np.abs(taus[b] - otau)) <= 1e-12


@pytest.mark.parametrize("m,n,nrhs", [(2, 2, 1), (5, 3, 1), (30, 8, 2), (128, 64, 1), (512, 64, 1), (700, 40, 1)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gels_matches_oracle(oracle, m, n, nrhs, dtype):
    k = 4
    r = np.random.default_rng(m + 7 * n)
    A = r.standard_normal((k, m, n))
    B = r.standard_normal((k, nrhs, m))
    if dtype == torch.float32:
        A, B = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    tA = dev(api.from_matrices(torch.from_numpy(A)).numpy(), dtype)
    tB = dev(B, dtype)
    info = api.gels_batched(tA, tB)
    X = tB.cpu().double().numpy()[:, :, :n]
    assert info.cpu().tolist() == [0] * k
    for b in range(k):
        for c in range(nrhs):
            ref = oracle.gels(A[b], B[b, c]).ravel()
            cond = np.linalg.cond(A[b])
            tol = (1e-4 if dtype == torch.float32 else 1e-12) * max(cond, 1.0) * 10
            assert np.linalg.norm(X[b, c] - ref) <= tol * max(np.linalg.norm(ref), 1.0), (m, n, b, c)


def test_trsm_and_ormqr_multi_column(oracle):
    r = np.random.default_rng(3)
    k, m, n = 5, 40, 12
    A = r.standard_normal((k, m, n))
    C = r.standard_normal((k, 3, m))
    tA = dev(api.from_matrices(torch.from_numpy(A)).numpy())
    tau = api.geqrf_batched(tA)
    tC = dev(C)
    api.ormqr_batched(tA, tau, tC, trans=True)
    got = tC.cpu().numpy()
    for b in range(k):
        oqr, otau = oracle.geqrf(A[b])
        ref = oracle.ormqr(oqr, otau, C[b].T, trans=True)
        assert np.allclose(got[b].T, ref, atol=1e-12)


def _ragged_check(oracle, refine_steps, sample=300):
    from gputils_amd import workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged()
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    tA, tB = dev(Af), dev(Bf)
    plan(tA, tB, refine_steps=refine_steps)
    out = tB.cpu().numpy()
    assert np.array_equal(tA.cpu().numpy(), Af), "A must be read only"
    rng = np.random.default_rng(0)
    errs = []
    for s in rng.choice(len(ms), sample, replace=False):
        m, n = int(ms[s]), int(ns[s])
        A = Af[offA[s]:offA[s] + m * n].reshape(n, m).T
        b = Bf[offB[s]:offB[s] + m]
        ref = oracle.gels(A, b).ravel()
        x = out[offB[s]:offB[s] + n]
        # entries n..m-1 of each right-hand side are left untouched
        assert np.array_equal(out[offB[s] + n:offB[s] + m], b[n:]), s
        errs.append((np.linalg.norm(x - ref) / np.linalg.norm(ref), np.linalg.cond(A), m, n, int(s)))
    return errs


def test_ragged_least_squares_config(oracle):
    """4096 uneven tall inconsistent systems (n in [8,64], m in
    [max(16,n),512], b random): FP32 QR with FP64 norms, then two FP64
    augmented-system refinements, vs the FP64 oracle dgels.  The bound is the
    FP64 forward-error level, 1e-12 * cond(A) (times a safety factor 100)."""
    errs = _ragged_check(oracle, 2)
    for err, cond, m, n, s in errs:
        assert err <= 1e-10 * max(cond, 1.0), (s, m, n, err, cond)


def test_ragged_least_squares_one_refinement_step(oracle):
    """The configuration's method: FP32 QR with FP64 norms and one FP64
    augmented-system refinement step.  One step contracts the FP32-level
    error by about u32 * cond, so the forward error is bounded by
    1e-9 * cond(A) (measured maximum 4e-12 * cond over 400 systems)."""
    errs = _ragged_check(oracle, 1)
    for err, cond, m, n, s in errs:
        assert err <= 1e-9 * max(cond, 1.0), (s, m, n, err, cond)


def test_ragged_least_squares_unrefined_is_fp32_level(oracle):
    """Without refinement the result carries the FP32 least-squares error
    (~u32 * cond^2 for a large residual); this fixes that the refinement, not
    the factorisation, supplies the FP64 accuracy above."""
    errs = _ragged_check(oracle, 0, sample=100)
    worst = max(e / max(c, 1.0) ** 2 for e, c, *_ in errs)
    assert worst <= 1e-4
    assert max(e for e, *_ in errs) > 1e-9


def _ragged_solve(plan, tA, Bf, steps):
    tB = dev(Bf)
    plan(tA, tB, refine_steps=steps)
    return tB.cpu().numpy()


def _bucket_sizes(ms, ns):
    """Systems per LDS size class (FP32 m*n <= 2048 .. 32768), largest first."""
    e = ms.astype(np.int64) * ns
    out = {}
    for x in e:
        cls = next(c for c in (2048, 4096, 8192, 16384, 32768) if x <= c)
        out[cls] = out.get(cls, 0) + 1
    return [(c, out[c]) for c in sorted(out, reverse=True)]


def test_ragged_plan_buckets_and_default_shape():
    from gputils_amd import workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged()
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    desc = plan.describe()
    assert [(d["maxe"], d["systems"]) for d in desc] == _bucket_sizes(ms, ns)
    for d in desc:
        assert d["method"] == 0 and d["waves"] == 8 and d["queue"] == 0
        assert d["per_cu"] >= 1 and 1 <= d["grid"] <= d["systems"]


_SHAPES = {0: (2, 4, 8, 16), 1: (8, 16)}


@pytest.mark.parametrize("method", [0, 1])
def test_ragged_every_launch_shape_gives_identical_results(method):
    """For one factorisation method the launch shape (wavefronts per
    workgroup, workgroups per CU, static ranges or the work queue, and so
    the assignment of systems to workgroups) changes no system's
    arithmetic: every shape reproduces the method's first shape bit for
    bit, also where a workgroup solves several systems in turn (per_cu = 1),
    and repeated queue launches (the counters reset themselves) agree."""
    from gputils_amd import workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged()
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    tA = dev(Af)
    nb = len(plan.describe())
    ref = None
    for queue in (0, 1):
        for waves in _SHAPES[method]:
            for per_cu in (1, 2, 3):
                for b in range(nb):
                    plan.configure(b, method, waves, per_cu, queue)
                for d in plan.describe():
                    assert (d["method"], d["waves"], d["per_cu"], d["queue"]) == (method, waves, per_cu, queue)
                    assert 1 <= d["grid"] <= d["systems"]
                got = _ragged_solve(plan, tA, Bf, 1)
                if ref is None:
                    ref = got
                assert np.array_equal(got, ref), (queue, waves, per_cu)
                if queue:
                    assert np.array_equal(_ragged_solve(plan, tA, Bf, 1), ref), ("repeat", waves, per_cu)


def test_ragged_register_panel_method_matches_oracle(oracle):
    """Method 1 (register panels) to the same bar as the default method:
    one and two refinement steps against the FP64 oracle dgels."""
    from gputils_amd import workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged()
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    for b in range(len(plan.describe())):
        plan.configure(b, 1, 16, 1)
    tA = dev(Af)
    rng = np.random.default_rng(0)
    sample = rng.choice(len(ms), 200, replace=False)
    for steps, bound in ((1, 1e-9), (2, 1e-10)):
        out = _ragged_solve(plan, tA, Bf, steps)
        for s in sample:
            m, n = int(ms[s]), int(ns[s])
            A = Af[offA[s]:offA[s] + m * n].reshape(n, m).T
            b = Bf[offB[s]:offB[s] + m]
            ref = oracle.gels(A, b).ravel()
            x = out[offB[s]:offB[s] + n]
            assert np.array_equal(out[offB[s] + n:offB[s] + m], b[n:]), s
            err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
            assert err <= bound * max(np.linalg.cond(A), 1.0), (steps, s, m, n, err)


def test_ragged_autotune_keeps_b_and_results():
    """The autotuner times the shapes on a copy of B (B is not written) and
    leaves a valid shape per bucket; the tuned plan solves bit-identically
    to a plan given the same shapes by hand."""
    from gputils_amd import workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged()
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    tA, tB = dev(Af), dev(Bf)
    desc = plan.tune(tA, tB, refine_steps=1)
    assert np.array_equal(tB.cpu().numpy(), Bf)
    assert [(d["maxe"], d["systems"]) for d in desc] == _bucket_sizes(ms, ns)
    for d in desc:
        assert d["waves"] in _SHAPES[d["method"]] and d["per_cu"] >= 1 and 1 <= d["grid"] <= d["systems"]
        assert d["queue"] in (0, 1)
    tuned = _ragged_solve(plan, tA, Bf, 1)
    other = api.RaggedLeastSquares(ms, ns, offA, offB)
    for b, d in enumerate(desc):
        other.configure(b, d["method"], d["waves"], d["per_cu"], d["queue"])
    assert np.array_equal(_ragged_solve(other, tA, Bf, 1), tuned)


def test_ragged_concurrent_and_sequential_buckets_agree():
    """The buckets write disjoint systems: running them concurrently on the
    plan's streams or one after another gives bit-identical results, and
    the caller's stream is ordered after either."""
    from gputils_amd import workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged()
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    assert plan.concurrent
    tA = dev(Af)
    conc = _ragged_solve(plan, tA, Bf, 1)
    plan.concurrent = False
    assert not plan.concurrent
    seq = _ragged_solve(plan, tA, Bf, 1)
    assert np.array_equal(conc, seq)
    # ordering on the caller's stream: the solve, then a read of B queued
    # behind it, with no host synchronisation in between
    plan.concurrent = True
    tB = dev(Bf)
    plan(tA, tB, refine_steps=1)
    snap = tB.clone()
    assert np.array_equal(snap.cpu().numpy(), conc)


@pytest.mark.parametrize("queue,concurrent", [(0, False), (1, False), (0, True), (1, True)])
def test_ragged_plan_graph_replays_match_eager(queue, concurrent):
    """One plan execution captured in a HIP graph (the concurrent buckets
    fork from and join back to the capturing stream by events; the work
    queue's counters reset themselves, no memset node) and replayed with
    eager work between replays: every replay equals the eager solve."""
    from gputils_amd import workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged(seed=9, count=1024)
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    for b in range(len(plan.describe())):
        plan.configure(b, 0, 8, 2, queue)
    plan.concurrent = concurrent
    tA = dev(Af)
    eager = _ragged_solve(plan, tA, Bf, 1)
    B0 = dev(Bf)
    tB = B0.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        plan(tA, tB, refine_steps=1)  # warm-up outside the capture
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            plan(tA, tB, refine_steps=1)
    torch.cuda.synchronize()
    for rep in range(4):
        tB.copy_(B0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(tB.cpu().numpy(), eager), rep
        _ = (tA[:4096] * 2).sum().item()  # eager work between replays


# edge shapes of the ragged plan: one column, one row, square systems, the
# LDS bucket boundaries m n = 2048 / 4096 / 8192 / 16384 on both sides, the
# largest system (512 x 64 = 32768, the last bucket's capacity) and a
# consistent system (b in the range of A, zero residual)
_EDGE_SHAPES = [(1, 1), (2, 1), (2, 2), (5, 5), (16, 1), (512, 1), (64, 64), (33, 32),
                (512, 4), (455, 5), (512, 8), (293, 14), (512, 16), (257, 32), (512, 32),
                (512, 33), (512, 64), (511, 64), (100, 64), (300, 7)]


@pytest.mark.parametrize("method,waves", [(0, 2), (0, 16), (1, 8), (1, 16)])
def test_ragged_edge_shapes_match_oracle(oracle, method, waves):
    """Systems at the plan's size limits and bucket boundaries, one and two
    refinement steps, against the FP64 oracle dgels, to the bounds of the
    configuration tests (1e-9 and 1e-10 times cond(A)).  The last system is
    consistent (b = A x): its solution must come out at the FP64 level
    too.  Entries n..m-1 of b are left as they were."""
    rng = np.random.default_rng(11)
    shapes = _EDGE_SHAPES
    ms = np.array([m for m, _ in shapes], dtype=np.int32)
    ns = np.array([n for _, n in shapes], dtype=np.int32)
    offA = np.concatenate([[0], np.cumsum(ms.astype(np.int64) * ns)[:-1]])
    offB = np.concatenate([[0], np.cumsum(ms.astype(np.int64))[:-1]])
    Af = rng.standard_normal(int((ms.astype(np.int64) * ns).sum()))
    Bf = rng.standard_normal(int(ms.sum()))
    # consistent last system: b = A x_true
    m, n = shapes[-1]
    A_last = Af[offA[-1]:offA[-1] + m * n].reshape(n, m).T
    Bf[offB[-1]:offB[-1] + m] = A_last @ rng.standard_normal(n)
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    for b in range(len(plan.describe())):
        plan.configure(b, method, waves, 1)
    tA = dev(Af)
    for steps, bound in ((1, 1e-9), (2, 1e-10)):
        out = _ragged_solve(plan, tA, Bf, steps)
        for s, (m, n) in enumerate(shapes):
            A = Af[offA[s]:offA[s] + m * n].reshape(n, m).T
            b = Bf[offB[s]:offB[s] + m]
            ref = oracle.gels(A, b).ravel()
            x = out[offB[s]:offB[s] + n]
            assert np.array_equal(out[offB[s] + n:offB[s] + m], b[n:]), (s, m, n)
            err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
            assert err <= bound * max(np.linalg.cond(A), 1.0), (steps, s, m, n, err)


def test_ragged_plan_configure_rejects_bad_shapes():
    from gputils_amd import workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged(count=64)
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    with pytest.raises(api.GputilsError):
        plan.configure(0, 2, 8, 1)    # method must be 0 or 1
    with pytest.raises(api.GputilsError):
        plan.configure(0, 1, 4, 1)    # method 1 takes 8 or 16 wavefronts
    with pytest.raises(api.GputilsError):
        plan.configure(0, 0, 3, 1)    # method 0 takes 2, 4, 8 or 16
    with pytest.raises(api.GputilsError):
        plan.configure(0, 0, 4, 0)    # per_cu >= 1
    with pytest.raises(api.GputilsError):
        plan.configure(len(plan.describe()), 0, 4, 1)  # no such bucket
    with pytest.raises(api.GputilsError):
        plan.configure(0, 0, 4, 1, 2)  # queue must be 0 or 1


def test_ragged_plan_rejects_bad_sizes():
    with pytest.raises(api.GputilsError):
        api.RaggedLeastSquares([10], [20], [0], [0])   # n > m
    with pytest.raises(api.GputilsError):
        api.RaggedLeastSquares([600], [10], [0], [0])  # m > 512
