"""GPU parity of the heterogeneous dwell batch engine (uneven P, R and
templates in one run) against the FP64 oracle, and invariance of each dwell's
result to the batch it runs in, the launch form (graph or not) and tuning."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

GUARD, REF, PFA = (2, 2), (8, 8), 1e-6
# (P, R, T): power-of-two R (one circular FFT per pulse) and R = 256*k with
# k not a power of two (overlap-save segments), P from 16 to 256; the last two
# have range lengths that are not a multiple of any strip chunk (partial last
# chunk and strip in the strip Doppler+CFAR jobs, P = 64 and P = 256)
MIX = [(32, 1024, 64), (64, 1280, 200), (128, 2048, 512), (16, 3328, 1000), (256, 1024, 100),
       (128, 4096, 256), (64, 1792, 64), (32, 16384 - 256, 1024), (64, 1000, 100), (256, 2000, 300)]


@pytest.fixture(scope="module")
def setup():
    from zkx_dev_amd import engine, synth
    eng = engine.DwellEngine(pc_window="hamming", normalize=True, doppler_window="hann", guard=GUARD,
                             reference=REF, pfa=PFA, lanes=2)
    dwells, tids = [], []
    for i, (p, r, t) in enumerate(MIX):
        d = synth.make_dwell(p, r, t, n_targets=3, snr_db=-10, seed=500 + i)
        dwells.append(d)
        tids.append(eng.add_template(torch.from_numpy(d.template)))
    descs, total = engine.make_descs([(p, r) for p, r, _ in MIX], tids)
    x = torch.from_numpy(np.concatenate([d.x.reshape(-1) for d in dwells])).cuda()
    return eng, engine, dwells, descs, total, x


def _outputs(total, rd=True):
    thr = torch.empty(total, dtype=torch.float32, device="cuda")
    det = torch.empty(total, dtype=torch.uint8, device="cuda")
    m = torch.empty(total, dtype=torch.complex64, device="cuda") if rd else None
    return thr, det, m


def test_plan_uses_overlap_save_for_non_pow2(setup):
    eng, _, _, descs, _, _ = setup
    n, s = eng.plan(descs)
    for (p, r, t), nn, ss in zip(MIX, n, s):
        if r & (r - 1) == 0:
            assert nn == r and ss == 1
        else:
            assert nn > t and nn & (nn - 1) == 0 and ss * (min(nn - t + 1, r)) >= r


def test_mixed_batch_vs_fp64(setup):
    eng, _, dwells, descs, total, x = setup
    thr, det, rd = _outputs(total)
    eng.run(x, descs, thr, det, rd)
    torch.cuda.synchronize()
    thr, det, rd = thr.cpu().numpy(), det.cpu().numpy(), rd.cpu().numpy()
    for (p, r, t), d, row in zip(MIX, dwells, descs):
        o = int(row["output_offset"])
        sl = slice(o, o + p * r)
        print(f"dwell P={p} R={r} T={t}", flush=True)
        ref = orc.pulse_doppler_fft64(orc.pulse_compression_fft64(d.x, d.template, window="hamming",
                                                                  normalize=True), window="hann")
        m = rd[sl].reshape(p, r)
        assert orc.rel_l2(m, ref) < 1e-5
        pw = (m.real.astype(np.float64) ** 2 + m.imag.astype(np.float64) ** 2).astype(np.float32)
        exact = orc.ca_cfar_exact(pw, GUARD, REF, PFA)
        t_gpu = thr[sl].reshape(p, r)
        assert np.allclose(t_gpu, exact, rtol=1e-5, atol=1e-30)
        near = np.abs(pw.astype(np.float64) - exact) <= 1e-6 * np.abs(exact)
        bad = (det[sl].reshape(p, r) != (pw.astype(np.float64) > exact)) & ~near
        assert not bad.any(), f"{int(bad.sum())} mask cells differ (near-threshold excused: {int(near.sum())})"
        hits = {(int(a), int(b)) for a, b in zip(*np.nonzero(det[sl].reshape(p, r)))}
        if p > 2 * (GUARD[0] + REF[0]):
            assert any((f, rb) in hits for rb, f in zip(d.range_bins, d.doppler_bins))


def test_each_dwell_same_alone_and_in_mix(setup):
    eng, engine, dwells, descs, total, x = setup
    thr, det, rd = _outputs(total)
    eng.run(x, descs, thr, det, rd)
    for i, row in enumerate(descs):
        p, r = int(row["num_pulses"]), int(row["range_bins"])
        o = int(row["output_offset"])
        one = descs[i:i + 1].copy()
        one["output_offset"] = 0
        t1, d1, m1 = _outputs(p * r)
        eng.run(x, one, t1, d1, m1, lane=1)
        torch.cuda.synchronize()
        assert torch.equal(t1, thr[o:o + p * r])
        assert torch.equal(d1, det[o:o + p * r])
        assert torch.equal(m1.view(torch.float32), rd[o:o + p * r].view(torch.float32))


def test_graph_and_direct_launch_agree(setup):
    eng, _, _, descs, total, x = setup
    a = _outputs(total)
    b = _outputs(total)
    eng.set_options(use_graph=True)
    for _ in range(3):  # replays with eager work between them
        eng.run(x, descs, *a)
        torch.cuda.synchronize()
    eng.set_options(use_graph=False)
    eng.run(x, descs, *b)
    torch.cuda.synchronize()
    eng.set_options(use_graph=True)
    for u, v in zip(a, b):
        assert torch.equal(u.view(torch.uint8), v.view(torch.uint8))


def test_matches_single_shape_pipeline(setup):
    from zkx_dev_amd import radar
    _, _, dwells, descs, total, x = setup
    eng, engine = setup[0], setup[1]
    thr, det, rd = _outputs(total)
    eng.run(x, descs, thr, det, rd)
    torch.cuda.synchronize()
    for (p, r, t), d, row in zip(MIX, dwells, descs):
        if r & (r - 1) or p > 128:
            continue
        o = int(row["output_offset"])
        pipe = radar.DwellPipeline(d.template, p, r, guard=GUARD, reference=REF, pfa=PFA, cfar_mode=1)
        m = torch.empty((1, p, r), dtype=torch.complex64, device="cuda")
        th, de = pipe.run(torch.from_numpy(d.x).cuda(), rd_map=m)
        torch.cuda.synchronize()
        assert orc.rel_l2(rd[o:o + p * r].cpu().numpy(), m.reshape(-1).cpu().numpy()) < 1e-6
        agree = (de.reshape(-1) == det[o:o + p * r]).float().mean().item()
        assert agree > 0.9999


def test_autotune_and_tuning_roundtrip(setup):
    eng, engine, _, descs, total, x = setup
    thr, det, rd = _outputs(total)
    eng.run(x, descs, thr, det, rd)
    eng.autotune(x, descs, thr, det)
    text = eng.tuning()
    assert text.startswith("grid ")
    n1, s1 = eng.plan(descs)
    e2 = engine.DwellEngine(guard=GUARD, reference=REF, pfa=PFA)
    for t in eng._templates:
        e2.add_template(torch.ones(t, dtype=torch.complex64))
    e2.load_tuning(text)
    n2, s2 = e2.plan(descs)
    assert np.array_equal(n1, n2) and np.array_equal(s1, s2)
    # tuned results still meet the FP64 bar on a dwell that has OLS choices
    t2, d2, m2 = _outputs(total)
    eng.run(x, descs, t2, d2, m2)
    torch.cuda.synchronize()
    assert orc.rel_l2(m2.cpu().numpy(), rd.cpu().numpy()) < 1e-5
    e2.close()


def test_validation(setup):
    from zkx_dev_amd._lib import ZkxError
    eng, engine, _, descs, total, x = setup
    thr, det, _ = _outputs(total, rd=False)
    bad = descs.copy()
    bad[0]["num_pulses"] = 48
    with pytest.raises(ZkxError) as e:
        eng.run(x, bad, thr, det)
    assert e.value.code == 1002
    bad = descs.copy()
    bad[1]["template_id"] = 99
    with pytest.raises(ZkxError) as e:
        eng.run(x, bad, thr, det)
    assert e.value.code == 1001
    with pytest.raises(ValueError):
        eng.run(x[:10], descs, thr, det)


def test_small_grids_match(setup):
    """Few persistent blocks, so every block runs many jobs of different
    sizes back to back (LDS reuse between jobs) — results must not change."""
    eng, _, _, descs, total, x = setup
    a = _outputs(total)
    eng.run(x, descs, *a)
    torch.cuda.synchronize()
    eng.set_options(use_graph=True, pc_grid=8, dc_grid=8)
    b = _outputs(total)
    eng.run(x, descs, *b)
    torch.cuda.synchronize()
    eng.set_options(use_graph=True, pc_grid=512, dc_grid=256)
    for u, v in zip(a, b):
        assert torch.equal(u.view(torch.uint8), v.view(torch.uint8))


def test_engine_other_cfar_window():
    """A window other than the compiled-in one (runtime cells).  The engine
    sizes LDS for P up to 256, which needs a range halo of at least 10 bins."""
    from zkx_dev_amd import engine, synth
    guard, ref, pfa = (1, 3), (4, 8), 1e-5
    eng = engine.DwellEngine(guard=guard, reference=ref, pfa=pfa)
    mix = [(32, 1280, 100), (64, 1024, 64)]
    dwells = [synth.make_dwell(p, r, t, n_targets=2, seed=900 + i) for i, (p, r, t) in enumerate(mix)]
    tids = [eng.add_template(torch.from_numpy(d.template)) for d in dwells]
    descs, total = engine.make_descs([(p, r) for p, r, _ in mix], tids)
    x = torch.from_numpy(np.concatenate([d.x.reshape(-1) for d in dwells])).cuda()
    thr, det, rd = _outputs(total)
    eng.run(x, descs, thr, det, rd)
    torch.cuda.synchronize()
    for (p, r, t), d, row in zip(mix, dwells, descs):
        o = int(row["output_offset"])
        m = rd[o:o + p * r].cpu().numpy().reshape(p, r)
        pw = (m.real.astype(np.float64) ** 2 + m.imag.astype(np.float64) ** 2).astype(np.float32)
        exact = orc.ca_cfar_exact(pw, guard, ref, pfa)
        assert np.allclose(thr[o:o + p * r].cpu().numpy().reshape(p, r), exact, rtol=1e-5, atol=1e-30)
        near = np.abs(pw.astype(np.float64) - exact) <= 1e-6 * np.abs(exact)
        bad = (det[o:o + p * r].cpu().numpy().reshape(p, r) != (pw.astype(np.float64) > exact)) & ~near
        assert not bad.any()
    eng.close()
