"""GPU test of the learned warm start on a small parametric family: training
solves from the batched engine, POD bases and MLP fitted on the device, then
held-out parameters solved cold, from the POD mean alone, and from the
network's prediction. Every solve must reach the tolerance (the cold solve is
the fallback and the judge); the network must save iterations over cold
starts, and its gain over the mean-only start is reported."""
import numpy as np
import pytest

from nirnaya_amd.engine import MODE_PDLP, default_params
from nirnaya_amd.warmstart import ParametricFamily, Pod, WarmStart, solve_warm

pytestmark = pytest.mark.gpu


def test_warm_start_saves_iterations(engine):
    fam = ParametricFamily(20264, 400, 1000, 5)
    params = default_params(mode=MODE_PDLP, max_iters=400000, tol=1e-6)
    th = fam.thetas(192, 1)
    b = engine.batch([fam.instance(t) for t in th], params)
    try:
        b.solve()
        res = [b.result(i) for i in range(len(th))]
    finally:
        b.close()
    ok = [i for i, r in enumerate(res) if r.status == 0]
    assert len(ok) >= 180
    pod_x = Pod(engine, np.stack([res[i].x for i in ok]), 64)
    pod_y = Pod(engine, np.stack([res[i].y for i in ok]), 64)
    ws = WarmStart(engine, pod_x, pod_y, hidden=512, max_rows=256)
    losses = ws.fit(th[ok], epochs=300, batch=64, lr=1e-3, seed=3)
    print(f"ranks {pod_x.rank}/{pod_y.rank}; loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.5 * losses[0]
    cold, mean_only, warm = [], [], []
    for t in fam.thetas(12, 2):
        lp = fam.instance(t)
        c = engine.solve(lp, params)
        m = engine.solve(lp, params, x0=pod_x.mean, y0=pod_y.mean)
        x0, y0 = ws.predict(t)
        w, rec = solve_warm(engine, lp, params, x0, y0)
        assert c.status == 0 and m.status == 0 and w.status == 0 and not rec.fallback
        cold.append(c.iterations)
        mean_only.append(m.iterations)
        warm.append(w.iterations)
    print(f"iterations cold {cold}\nmean-only {mean_only}\nwarm {warm}")
    print(f"median cold {np.median(cold)} mean-only {np.median(mean_only)} warm {np.median(warm)}")
    assert np.median(warm) < np.median(cold)
    ws.net.close()
