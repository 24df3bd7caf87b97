"""GPU tests of the continuous-batching server: each request's result is the
same alone and in a mixed batch (bit for bit), admission respects the HBM
budget, an open-loop Poisson replay completes, and bad requests are refused
at submit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

GUARD, REF, PFA = (2, 2), (8, 8), 1e-6
SHAPES = [(64, 2048), (128, 4096), (256, 8192)]
TAPS = [128, 256, 384, 512]


@pytest.fixture(scope="module")
def env():
    from zkx_dev_amd import engine, serve, synth
    specs = [(p, r, TAPS[i % len(TAPS)]) for i, (p, r) in enumerate(SHAPES * 2)]
    ub = synth.make_uneven_batch_device(specs, seed=77)
    eng = engine.DwellEngine(guard=GUARD, reference=REF, pfa=PFA, lanes=4)
    tids = [eng.add_template(torch.from_numpy(synth.lfm_chirp(t))) for t in TAPS]
    reqs = [(ub.offsets[i], p, r, tids[i % len(TAPS)]) for i, (p, r, _) in enumerate(specs)]
    return eng, serve, ub, reqs


def test_same_alone_and_in_mix(env):
    eng, serve, ub, reqs = env
    srv = serve.Server(eng, ub.x, budget_bytes=8 << 30, lanes=3, keep_outputs=True)
    alone = []
    for off, p, r, t in reqs:
        k = srv.submit(off, p, r, t)
        assert srv.wait(k, 10000)
        _, thr, det = srv.result(k, p, r)
        torch.cuda.synchronize()
        alone.append((thr.clone(), det.clone()))
        srv.release(k)
    b0 = srv.stats()["batches"]
    srv.hold(True)
    tickets = [srv.submit(*q) for q in reqs]  # queued together, launched as one batch
    srv.hold(False)
    for k in tickets:
        assert srv.wait(k, 10000)
    st = srv.stats()
    assert st["batches"] - b0 == 1, st
    for k, (off, p, r, t), (a_thr, a_det) in zip(tickets, reqs, alone):
        lat, thr, det = srv.result(k, p, r)
        torch.cuda.synchronize()
        assert lat > 0
        assert torch.equal(thr.view(torch.int32), a_thr.view(torch.int32))
        assert torch.equal(det, a_det)
        srv.release(k)
    assert srv.stats()["reserved_bytes"] == 0
    srv.close()


def test_results_match_engine_and_oracle(env):
    eng, serve, ub, reqs = env
    srv = serve.Server(eng, ub.x, budget_bytes=8 << 30, lanes=2, keep_outputs=True)
    off, p, r, t = reqs[0]
    k = srv.submit(off, p, r, t)
    assert srv.wait(k, 10000)
    _, thr, det = srv.result(k, p, r)
    torch.cuda.synchronize()
    from zkx_dev_amd import engine as E
    d, total = E.make_descs([(p, r)], [t])
    d["input_offset"] = off
    t2 = torch.empty(total, dtype=torch.float32, device="cuda")
    d2 = torch.empty(total, dtype=torch.uint8, device="cuda")
    m2 = torch.empty(total, dtype=torch.complex64, device="cuda")
    eng.run(ub.x, d, t2, d2, m2, lane=3)
    torch.cuda.synchronize()
    assert torch.equal(thr.reshape(-1).view(torch.int32), t2.view(torch.int32))
    assert torch.equal(det.reshape(-1), d2)
    m = m2.cpu().numpy().reshape(p, r)
    pw = (m.real.astype(np.float64) ** 2 + m.imag.astype(np.float64) ** 2).astype(np.float32)
    exact = orc.ca_cfar_exact(pw, GUARD, REF, PFA)
    near = np.abs(pw.astype(np.float64) - exact) <= 1e-6 * np.abs(exact)
    bad = (det.cpu().numpy() != (pw.astype(np.float64) > exact)) & ~near
    assert not bad.any()
    srv.release(k)
    srv.close()


def test_budget_admission(env):
    eng, serve, ub, reqs = env
    big = max(p * r for _, p, r, _ in reqs)
    budget = int(13 * big * 1.5)  # room for one large request at a time
    srv = serve.Server(eng, ub.x, budget_bytes=budget, lanes=3, keep_outputs=False)
    tickets = [srv.submit(*q) for q in reqs * 2]
    for k in tickets:
        assert srv.wait(k, 20000)
    st = srv.stats()
    assert st["peak_reserved_bytes"] <= budget
    assert st["dwells_done"] == len(tickets)
    assert st["reserved_bytes"] == 0
    srv.close()


def test_poisson_replay(env):
    eng, serve, ub, reqs = env
    srv = serve.Server(eng, ub.x, budget_bytes=16 << 30, lanes=3, keep_outputs=False)
    rng = np.random.default_rng(3)
    n = 300
    arrivals = np.cumsum(rng.exponential(1e6 / 5000.0, n))  # 5k requests/s
    pick = rng.integers(0, len(reqs), n)
    lat = srv.replay(arrivals, [reqs[i][0] for i in pick], [reqs[i][1] for i in pick],
                     [reqs[i][2] for i in pick], [reqs[i][3] for i in pick])
    assert np.all(lat > 0) and np.all(np.isfinite(lat))
    print(f"p50 {np.percentile(lat, 50):.0f} us, p99 {np.percentile(lat, 99):.0f} us", flush=True)
    assert srv.stats()["dwells_done"] == n
    srv.close()


def test_submit_validation(env):
    from zkx_dev_amd._lib import ZkxError
    eng, serve, ub, reqs = env
    srv = serve.Server(eng, ub.x, budget_bytes=1 << 30, lanes=1)
    off, p, r, t = reqs[0]
    with pytest.raises(ZkxError) as e:
        srv.submit(off, 48, r, t)
    assert e.value.code == 1002
    with pytest.raises(ZkxError) as e:
        srv.submit(off, p, r, 99)
    assert e.value.code == 1001
    with pytest.raises(ZkxError) as e:
        srv.submit(ub.x.numel(), p, r, t)
    assert e.value.code == 1001
    small = serve.Server(eng, ub.x, budget_bytes=13 * 1024, lanes=1)
    with pytest.raises(ZkxError) as e:
        small.submit(off, p, r, t)
    assert e.value.code == 2003
    small.close()
    srv.close()
