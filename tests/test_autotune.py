"""The autotuner: its candidate set (CPU) and a tuning run on the GPU that
must return a shape whose image equals the default shape's, and must come
back from its cache file unchanged."""
import numpy as np
import pytest


def _has(cands, **kw):
    return any(all(c.get(k) == v for k, v in kw.items()) for c in cands)


def test_candidates_cover_the_planner_shapes_within_limits():
    from imaging2d_wlsq_amd import autotune
    wide = autotune.candidates(True, True, 8192)
    assert _has(wide, x_per_thread=4, shots_per_group=1, block=1024)
    assert _has(wide, x_per_thread=8, shots_per_group=1)
    assert not _has(wide, x_per_thread=2)  # 4096 strips: no 2-x shape fits
    narrow = autotune.candidates(True, True, 1024)
    assert _has(narrow, x_per_thread=2, shots_per_group=4)
    assert _has(narrow, x_per_thread=2, shots_per_group=2, tap_loads=2)
    assert _has(narrow, tables_symmetric=0) and _has(narrow, tables_symmetric=1)
    for c in wide + narrow:
        assert not (c["x_per_thread"] == 2 and c["shots_per_group"] == 1)
        assert not (c["x_per_thread"] == 8 and c["shots_per_group"] != 1)
    assert not _has(autotune.candidates(False, False, 1024), tables_symmetric=1)
    # two strips per thread where one would do: 1024 threads, one shot, 4 x
    assert _has(narrow, strips_per_thread=2, x_per_thread=4, shots_per_group=1, block=1024)
    assert not _has(wide, strips_per_thread=2)  # 8192 wide: the planner's own choice
    for c in narrow:
        if c.get("strips_per_thread") == 2:
            assert c["x_per_thread"] == 4 and c["shots_per_group"] == 1 and c["lds_budget"] > 0


@pytest.mark.gpu
def test_tuned_shape_gives_the_default_image_and_is_cached(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from imaging2d_wlsq_amd import autotune, batch, synth
    rng = np.random.default_rng(5)
    m = synth.gradient_model(rng, 30, 700, 10.0, 10.0, 1500.0, 3500.0, 2, vmin=1500.0, vmax=3500.0)
    t = synth.make_tables(m, 15, 1024, 2.0, 40.0)
    p = synth.Problem("tune", m, t, [])
    shots = [synth.ShotGeom(int(x0), int(nx), 30, nx / 2.0, 2.0 + 38.0 * np.arange(nf) / (nf - 1), k)
             for k, (x0, nx, nf) in enumerate([(0, 700, 9), (100, 400, 6), (50, 650, 12), (0, 300, 5)])]
    sv = batch.Survey(p)
    ref = sv.batch(shots)
    img_ref = ref.run().clone()
    cache = str(tmp_path / "tune.json")
    autotune._MEM.clear()
    b = sv.batch(shots, tune=True, reps=1)
    tuning = b.tuning
    print("tuned", tuning, flush=True)
    img = b.run().clone()
    torch.cuda.synchronize()
    err = float(torch.linalg.norm(img - img_ref) / torch.linalg.norm(img_ref))
    assert err <= 1e-5, err
    # the cache file round trip returns the same choice without timing
    args = (sv.slowness, b.omega, t.kappa0, t.dkappa, t.nk, sv.tf, sv.tb, b.F, b.B, b.image)
    autotune._MEM.clear()
    first = autotune.tune(b.specs, t.M, *args, sv.conj, sv.sym, reps=1, cache_file=cache)
    autotune._MEM.clear()
    again = autotune.tune(b.specs, t.M, *args, sv.conj, sv.sym, reps=1, cache_file=cache,
                          cands=[])  # no candidates: only the cache can answer
    assert again == first
    ref.close()
    b.close()


def test_longest_run_counts_shots_that_can_share_a_group():
    from imaging2d_wlsq_amd import autotune, device
    mk = lambda nx, x0: device.ShotSpec(0, 0, 0, x0, nx, 1000, 1000, nx, 50, 10, 0)  # noqa: E731
    assert autotune.longest_run([mk(100, 0)]) == 1
    assert autotune.longest_run([mk(100, 0), mk(100, 0), mk(200, 0), mk(100, 0)]) == 2
    assert autotune.longest_run([mk(100, 0), mk(100, 5)]) == 1          # other window
    assert autotune.longest_run([mk(64, 0)] * 9) == 9
