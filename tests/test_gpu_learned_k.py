"""Learned k on the GPU against the oracle (oracle/learned_k.py, liboracle).

Bar:
* features: within rtol 1e-6 of the oracle (integer counts and FP64
  distances rounded to float32; log2 and sqrt come from different libms);
* logits: the MFMA network (bfloat16 operands, float32 accumulation) within
  2e-3 of the oracle's bfloat16-rounded FP64 network, relative to the
  largest logit; the chosen bucket equal wherever the oracle's two best
  expected costs differ by more than 1%;
* gradients: every parameter block within 2e-2 relative (Frobenius) of the
  oracle's bfloat16 backward pass; one Adam step equal to the oracle's Adam
  applied to the GPU gradient within 1e-6;
* the mesh built with learned k equals the fixed-k mesh with restarts bit for
  bit (statuses, certified k, faces, normals; geometry within 1e-12), also
  for a deliberately bad model that predicts k = 16 everywhere;
* trained on one point set and applied to another, the model cuts the
  restarts of the fixed-k build by more than 3x; with restarts that clip
  again from the first neighbour, the captured batched step is faster than
  the fixed-k one.  (With restarts that resume the saved cell, the default,
  a restart costs only its longer KNN row and the two steps take about the
  same time on uniform seeds; the test prints both.)
"""
import numpy as np
import pytest

from oracle import learned_k as L
from oracle import oracle as O
from proteusgpu_amd import api
from proteusgpu_amd.generators import gaussian_mixture, uniform_points

pytestmark = pytest.mark.gpu


def rand_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(L.NPARAM) * scale).astype(np.float32)


@pytest.mark.parametrize("dim,n,boundary,kind", [(3, 20000, "reflective", "uniform"), (2, 10000, "periodic", "uniform"),
                                              (3, 20000, "reflective", "clustered"),
                                              (2, 10000, "periodic", "clustered")])
def test_features_match_oracle(ctx, dim, n, boundary, kind):
    """Clustered: the crowded grid cells' points are scanned through their
    sub-grids (pruned by the 8th distance), the oracle scans them all."""
    print(f"features dim={dim} n={n} {boundary} {kind}", flush=True)
    if kind == "uniform":
        pts = uniform_points(n, dim, 41)
    else:
        pts = np.ascontiguousarray(gaussian_mixture(n, 8, 0.003, 0.03, 0.3, 43)[:, :dim])
    m = api.Mesh(ctx, pts, boundary=boundary, buff=0.12 if boundary == "periodic" else 0.04, k=32)
    Fg = m.features()
    ref = O.Mesh(pts, boundary, buff=0.12 if boundary == "periodic" else 0.04, k=32, restart=True)
    S, perm = ref.get("sorted_points"), ref.get("permutation")
    Fo_sorted = O.kfeatures(S[:, :dim], m.N_grid)
    Fo = np.empty_like(Fo_sorted)
    Fo[perm] = Fo_sorted
    np.testing.assert_allclose(Fg, Fo, rtol=1e-6, atol=1e-7)


def test_predict_matches_oracle(ctx):
    rng = np.random.default_rng(8)
    X = np.abs(rng.standard_normal((5000, 16))).astype(np.float32)
    P = rand_params(9)
    km = api.KModel(ctx)
    km.params = P
    km.set_cost(2.8)
    cls, lg = km.predict(X, logits=True)
    zo, _ = L.forward(P, X, bf16_ops=True)
    scale = np.max(np.abs(zo))
    assert np.max(np.abs(lg - zo)) <= 2e-3 * scale
    # bucket choice: equal wherever the decision is not a near tie
    co = L.decide(zo, 2.8)
    p = L.softmax(zo)
    Cm = np.array([[L.chain_cost(L.BUCKETS[b], L.KREP[c], 2.8) for c in range(6)] for b in range(6)])
    cost = p @ Cm.T
    srt = np.sort(cost, axis=1)
    clear = (srt[:, 1] - srt[:, 0]) > 0.01 * srt[:, 0]
    assert clear.sum() > 0.8 * len(X)
    assert np.array_equal(cls[clear], co[clear])


def test_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(10)
    X = np.abs(rng.standard_normal((1000, 16))).astype(np.float32)
    y = rng.integers(0, 6, 1000).astype(np.int32)
    P = rand_params(11)
    km = api.KModel(ctx)
    km.params = P
    loss_g, g = km.grad(X, y)
    loss_o, go = L.loss_grad(P, X, y, bf16_ops=True)
    assert abs(loss_g - loss_o) <= 1e-4 * abs(loss_o)
    gd, od = L.unpack(g), L.unpack(go)
    for name, _ in L.SHAPES:
        err = np.linalg.norm(gd[name] - od[name]) / np.linalg.norm(od[name])
        print(f"grad {name}: rel err {err:.2e}", flush=True)
        assert err < 2e-2, name


def test_adam_step_matches_oracle(ctx):
    rng = np.random.default_rng(12)
    X = np.abs(rng.standard_normal((512, 16))).astype(np.float32)
    y = rng.integers(0, 6, 512).astype(np.int32)
    P = rand_params(13)
    km = api.KModel(ctx)
    km.params = P
    _, g = km.grad(X, y)
    km.train_step(X, y, 1e-3)
    P1 = km.params
    Po, _, _ = L.adam(P.astype(np.float64), g.astype(np.float64), np.zeros(L.NPARAM), np.zeros(L.NPARAM), 1,
                      1e-3)
    np.testing.assert_allclose(P1, Po, rtol=0, atol=1e-6)


def test_training_reduces_loss(ctx):
    pts = uniform_points(30000, 3, 14)
    m = api.Mesh(ctx, pts, boundary="reflective", k=32)
    km = api.KModel(ctx, seed=3)
    l0 = km.train_mesh(m, steps=1, batch=4096, lr=1e-9)
    l1 = km.train_mesh(m, steps=300, batch=4096, lr=3e-3)
    print(f"loss {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < 0.85 * l0


def _same_mesh(a, b):
    for key in ("status", "kcert", "face_counts", "face_ptr", "neighbor_cell", "face_normal", "seeds"):
        assert np.array_equal(a[key], b[key]), key
    np.testing.assert_allclose(a["volumes"], b["volumes"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(a["face_area"], b["face_area"], rtol=1e-12, atol=1e-300)


def bad_model(ctx):
    """A model that predicts bucket 0 (k = 16) for every seed."""
    km = api.KModel(ctx)
    P = np.zeros(L.NPARAM, dtype=np.float32)
    P[5632] = 30.0
    km.params = P
    return km


@pytest.mark.parametrize("dim,n,boundary", [(3, 30000, "reflective"), (2, 20000, "periodic")])
def test_learned_mesh_equals_fixed_k(ctx, dim, n, boundary):
    print(f"learned mesh dim={dim} n={n} {boundary}", flush=True)
    buff = 0.12 if boundary == "periodic" else 0.04
    train = uniform_points(n, dim, 15)
    pts = uniform_points(n, dim, 16)
    k = 32 if dim == 3 else 16
    km = api.KModel(ctx, seed=5)
    km.train_mesh(api.Mesh(ctx, train, boundary=boundary, buff=buff, k=k), steps=200, batch=4096, lr=3e-3)
    plain = api.Mesh(ctx, pts, boundary=boundary, buff=buff, k=k)
    learned = api.Mesh(ctx, pts, boundary=boundary, buff=buff, k=k, kmodel=km)
    assert learned.n_failed == 0 and plain.n_failed == 0
    assert learned.n_class.sum() == learned.n_seeds
    _same_mesh(learned.copy_out(), plain.copy_out())
    bad = api.Mesh(ctx, pts, boundary=boundary, buff=buff, k=k, kmodel=bad_model(ctx))
    assert bad.n_class[0] == bad.n_seeds and bad.n_restarted > 0
    _same_mesh(bad.copy_out(), plain.copy_out())
    # and against the oracle
    ref = O.Mesh(pts, boundary, buff=buff, k=k, restart=True)
    g = learned.copy_out()
    for key in ("status", "kcert", "face_counts", "neighbor_cell"):
        assert np.array_equal(g[key], ref.get(key)), key


def test_learned_k_cuts_work(ctx):
    """Trained on one uniform set, applied to another of the same kind: the
    host-driven build restarts far fewer seeds, and the captured batched step
    (one box, 200k seeds, Sod) is faster with the same results."""
    from proteusgpu_amd.generators import sod_state
    train = uniform_points(200000, 3, 17)
    pts = uniform_points(200000, 3, 18)
    km = api.KModel(ctx, seed=6)
    km.train_mesh(api.Mesh(ctx, train, boundary="reflective", k=32), steps=400, batch=4096, lr=3e-3)
    plain = api.Mesh(ctx, pts, boundary="reflective", k=32)
    learned = api.Mesh(ctx, pts, boundary="reflective", k=32, kmodel=km)
    print(f"build: plain rows {plain.row_entries} restarted {plain.n_restarted}; learned rows "
          f"{learned.row_entries} restarted {learned.n_restarted} classes {learned.n_class}", flush=True)
    assert learned.n_restarted < 0.35 * plain.n_restarted
    rho, vel, E = sod_state(pts, 1.4)
    kw = dict(boundary="reflective", k=32, gamma=1.4, cfl=0.3, flux="physical", move_mesh=True)
    res, ms = [], {}
    try:
        for resume in (False, True):
            api.set_clip_resume(resume)
            for model in (None, km):
                B = api.Batch(ctx, np.array([len(pts)]), pts, rho, vel, E, **kw)
                if model is not None:
                    B.set_kmodel(model)
                t = [B.step() for _ in range(6)]
                ms[(resume, model is not None)] = float(np.median(t[1:]))
                res.append(B.get())
                B.free()
    finally:
        api.set_clip_resume(True)
    print(f"batched step, restarts from scratch: fixed k {ms[(False, False)]:.3f} ms, learned k "
          f"{ms[(False, True)]:.3f} ms; resumed restarts: fixed k {ms[(True, False)]:.3f} ms, learned k "
          f"{ms[(True, True)]:.3f} ms", flush=True)
    for r in res[1:]:
        for a, b in zip(res[0], r):
            assert np.array_equal(a, b)
    assert ms[(False, True)] < 0.97 * ms[(False, False)]


@pytest.mark.parametrize("dim,n,boundary", [(3, 30000, "reflective"), (2, 20000, "periodic")])
def test_learned_simulation_equals_fixed_k(ctx, dim, n, boundary):
    """Moving-mesh steps with learned k (from the second rebuild on, the
    seeds are clipped in the order of their previous certified k) give the
    fixed-k simulation bit for bit."""
    from proteusgpu_amd.generators import sod_state
    print(f"learned sim dim={dim} n={n} {boundary}", flush=True)
    buff = 0.12 if boundary == "periodic" else 0.04
    k = 32 if dim == 3 else 16
    pts = uniform_points(n, dim, 17)
    rho, vel, E = sod_state(pts, 1.4)
    kw = dict(boundary=boundary, buff=buff, k=k, gamma=1.4, cfl=0.3, flux="physical", move_mesh=True)
    km = api.KModel(ctx, seed=6)
    km.train_mesh(api.Mesh(ctx, uniform_points(n, dim, 18), boundary=boundary, buff=buff, k=k), steps=200,
                  batch=4096, lr=3e-3)
    runs = []
    for model in (None, km):
        sim = api.Simulation(ctx, pts, rho, vel, E, **kw)
        if model is not None:
            sim.set_kmodel(model)
        dts = [sim.step() for _ in range(4)]
        runs.append((dts, sim.get()))
        sim.free()
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)
