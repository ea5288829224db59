"""Learned initial iterate on the GPU: MFMA inference against a float64 restatement
of the same BF16 network, training gradients against torch autograd (float64, CPU),
the Adam step, and the residual judge (the start never changes the converged answer).

The MLP has no counterpart in the reference (its Newton loop starts from the
previous solution, StabFEM2.cpp:593-832); the checks here pin the kernels to the
mathematics of the network as documented in include/sfem.h."""
import numpy as np
import pytest
import torch

from cfdsfemmpi_amd import deck, meshgen, sfem

pytestmark = pytest.mark.gpu

NP = sfem.LEARNER_NPARAM
SL = {"W1": (0, (64, 12)), "b1": (768, (64,)), "W2": (832, (64, 64)), "b2": (4928, (64,)),
      "W3": (4992, (3, 64)), "b3": (5184, (3,)), "tau": (5187, (3,))}
TAU = np.array([0.5, 0.7, 1.3], np.float32)


def _tiny_cylinder(perturb=0.01, seed=1):
    return meshgen.cylinder_channel(nr=8, nth=32, n_left=4, n_right=12, n_bottom=4, n_top=4,
                                    perturb=perturb, seed=seed)


def _ctx(mesh, mu=0.01, reorder=1, rtol=1e-10):
    ctl = deck.Control(np.array([1.0, mu, 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0]), 2, 0.0, 0.1, 1e9, 100, 1, 1e-8)
    fem = sfem.StabFEM(device=0)
    fem.setMesh(mesh)
    fem.setControl(ctl)
    fem.prepareInputData()
    opts = sfem.solver_opts("bcgs", "ilu", rtol, 20000, reorder=reorder)
    fem.setSolver(opts)
    return fem, fem.ctx, opts


def _split(w):
    return {k: np.asarray(w[o:o + int(np.prod(s))], np.float64).reshape(s) for k, (o, s) in SL.items()}


def _bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _neighbours(mesh):
    nb = [set() for _ in range(mesh.nNode)]
    for e in mesh.conn:
        for a in e:
            nb[a].update(int(b) for b in e if b != a)
    return [np.array(sorted(s), np.int64) for s in nb]


def _activity(x32):
    """Divide float32 features by the node activity scale lambda (returned, float32)."""
    lam = np.sqrt((x32[:, :6].astype(np.float32) ** 2).sum(axis=1, dtype=np.float32) / np.float32(6)) \
        + np.float32(1e-3)
    return (x32 * (np.float32(1) / lam)[:, None]).astype(np.float32), lam.astype(np.float32)


def _features(S0, S1, S2, nbrs, isig):
    """The 12 inputs per node (float64 differences, scaled, rounded to float32, divided by
    the node's activity scale lambda), d1, d2 and lambda."""
    s0, s1, s2 = (np.asarray(S, np.float64).reshape(-1, 3) for S in (S0, S1, S2))
    d1, d2 = s0 - s1, s1 - s2
    m1 = np.array([d1[j].mean(axis=0) for j in nbrs])
    m2 = np.array([d2[j].mean(axis=0) for j in nbrs])
    x = np.concatenate([d1 * isig, d2 * isig, (m1 - d1) * isig, (m2 - d2) * isig], axis=1)
    x, lam = _activity(x.astype(np.float32))
    return x.astype(np.float64), d1, d2, lam.astype(np.float64)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _random_weights(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    w = (scale * rng.standard_normal(NP)).astype(np.float32)
    w[SL["tau"][0]:] = TAU
    return w


def _set_history(ctx, S0, S1, S2):
    """soln = S0, solnPrev = S1, solnPrev2 = S2 through the time-update shift."""
    ctx.reset_state()
    ctx.set_solution(S2)
    ctx.time_update(0.0)
    ctx.set_solution(S1)
    ctx.time_update(0.0)
    ctx.set_solution(S0)


def _smooth_states(mesh, seed):
    rng = np.random.default_rng(seed)
    xy = mesh.coords
    out = []
    base = np.stack([np.sin(0.3 * xy[:, 0]) * np.cos(0.5 * xy[:, 1]), np.cos(0.2 * xy[:, 0] + xy[:, 1]),
                     0.1 * xy[:, 0]], axis=1)
    vel = rng.standard_normal((1, 3)) * 0.05 + 0.02 * np.stack([np.cos(xy[:, 1]), np.sin(xy[:, 0]), xy[:, 1] / 10], 1)
    for k in range(3):
        out.append((base - k * vel - 0.3 * k * k * vel ** 2).ravel())
    return out  # S0, S1, S2


def test_quadratic_predictor_start():
    """SFEM_PRED_QUADRATIC starts each step from 3 s_n - 3 s_(n-1) + s_(n-2)."""
    mesh = _tiny_cylinder()
    fem, ctx, _ = _ctx(mesh)
    S0, S1, S2 = _smooth_states(mesh, 4)
    _set_history(ctx, S0, S1, S2)
    ctx.set_predictor(sfem.PRED_QUADRATIC)
    ctx.time_update(0.5)
    assert np.allclose(ctx.get_solution(), 3 * S0 - 3 * S1 + S2, rtol=0, atol=1e-14 * np.max(np.abs(S0)))


def test_zero_network_gives_the_quadratic_extrapolation():
    mesh = _tiny_cylinder()
    fem, ctx, _ = _ctx(mesh)
    L = sfem.Learner(ctx, 8, seed=3)
    L.set_weights(np.zeros(NP, np.float32))
    ctx.set_learner(L)
    S0, S1, S2 = _smooth_states(mesh, 0)
    _set_history(ctx, S0, S1, S2)
    cand = ctx.learned_candidate()
    assert np.allclose(cand, 3 * S0 - 3 * S1 + S2, rtol=0, atol=1e-14 * np.max(np.abs(S0)))


@pytest.mark.parametrize("ncomp", [2, 3])
def test_mfma_inference_matches_bf16_network(ncomp):
    """The MFMA forward pass equals the network with BF16-rounded weights, inputs and
    hidden activations (FP32 biases and accumulation) to accumulation-order accuracy;
    a transposed or misplaced operand tile would show O(1) differences.  With two
    corrected components the pressure keeps the quadratic extrapolation."""
    mesh = _tiny_cylinder()
    fem, ctx, _ = _ctx(mesh)
    L = sfem.Learner(ctx, 8, seed=5)
    with pytest.raises(sfem.SfemError):
        L.set_components(1)
    L.set_components(ncomp)
    w = _random_weights(11)
    L.set_weights(w)
    ctx.set_learner(L)
    S0, S1, S2 = _smooth_states(mesh, 1)
    _set_history(ctx, S0, S1, S2)
    cand = ctx.learned_candidate().reshape(-1, 3)
    d1 = (S0 - S1).reshape(-1, 3)
    sig = np.sqrt(np.mean(d1 ** 2, axis=0)).astype(np.float32).astype(np.float64)
    isig = (np.float32(1.0) / sig.astype(np.float32)).astype(np.float64)
    x, d1, d2, lam = _features(S0, S1, S2, _neighbours(mesh), isig)
    P = _split(w)
    h1 = _bf16(_silu(_bf16(x) @ _bf16(P["W1"]).T + P["b1"]))
    h2 = _bf16(_silu(h1 @ _bf16(P["W2"]).T + P["b2"]))
    y = h2 @ _bf16(P["W3"]).T + P["b3"]
    quad = S0.reshape(-1, 3) + 2 * d1 - d2
    if ncomp == 2:  # the pressure keeps the quadratic extrapolation
        assert np.allclose(cand[:, 2], quad[:, 2], rtol=0, atol=1e-14 * np.max(np.abs(S0)))
    y, y_gpu = y[:, :ncomp], ((cand - quad) / (sig * P["tau"] * lam[:, None]))[:, :ncomp]
    err = np.max(np.abs(y_gpu - y))
    assert np.max(np.abs(y)) > 0.5
    assert err <= 2e-2 * np.max(np.abs(y)), err
    # every output row checked, not only a few: the rows of one wavefront share nothing
    assert np.mean(np.abs(y_gpu - y)) <= 2e-3 * np.max(np.abs(y))


def _record_random_trajectory(ctx, L, mesh, nsnap, seed):
    rng = np.random.default_rng(seed)
    xy = mesh.coords
    snaps = []
    for t in range(nsnap):
        s = np.stack([np.sin(0.2 * xy[:, 0] + 0.3 * t), np.cos(0.3 * xy[:, 1] - 0.2 * t), 0.05 * t * xy[:, 0]], 1)
        s = s + 0.01 * rng.standard_normal(s.shape)
        ctx.set_solution(s.ravel())
        L.record(ctx, new_trajectory=(t == 0))
        snaps.append(s.ravel().astype(np.float32))
    return snaps


def _torch_loss_grad(w, snaps, windows, nodes, nbrs, mask, ncomp=2):
    """Mean squared error of the FP32 training network, float64 autograd on the CPU."""
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in _split(w).items()}
    X, T, M = [], [], []
    for t, i in zip(windows, nodes):
        s0, s1, s2, s3 = (snaps[t + d].astype(np.float64).reshape(-1, 3) for d in (0, -1, -2, 1))
        dd = s0 - s1
        rms = np.sqrt(np.mean(dd ** 2, axis=0)).astype(np.float32)
        isig = (np.float32(1.0) / rms).astype(np.float64)
        d1 = dd[i]
        d2 = s1[i] - s2[i]
        m1 = dd[nbrs[i]].mean(axis=0)
        m2 = (s1 - s2)[nbrs[i]].mean(axis=0)
        x = np.concatenate([d1 * isig, d2 * isig, (m1 - d1) * isig, (m2 - d2) * isig]).astype(np.float32)
        x, lam = _activity(x[None, :])
        X.append(x[0])
        T.append(((s3[i] - s0[i] - 2 * d1 + d2) * isig).astype(np.float32) / lam[0] / TAU)
        M.append(mask[3 * i:3 * i + 3] * np.array([1.0, 1.0, 1.0 if ncomp == 3 else 0.0]))
    X = torch.tensor(np.array(X, np.float64))
    T = torch.tensor(np.array(T, np.float64))
    M = torch.tensor(np.array(M, np.float64))
    h1 = torch.nn.functional.silu(X @ P["W1"].T + P["b1"])
    h2 = torch.nn.functional.silu(h1 @ P["W2"].T + P["b2"])
    y = h2 @ P["W3"].T + P["b3"]
    loss = (((y - T) * M) ** 2).sum() / (3 * len(windows))
    loss.backward()
    g = np.zeros(NP)
    for k, (o, s) in SL.items():
        if k != "tau":  # fixed, not trained: its gradient slot is zero
            g[o:o + int(np.prod(s))] = P[k].grad.numpy().ravel()
    return float(loss.detach()), g


def _free_mask(mesh):
    m = np.ones(mesh.nNode * 3)
    m[np.asarray(mesh.dbc_node, np.int64) * 3 + np.asarray(mesh.dbc_dof, np.int64)] = 0.0
    return m


@pytest.mark.parametrize("ncomp", [2, 3])
def test_training_gradient_matches_autograd(ncomp):
    """Loss and gradient of an explicit sample list against float64 autograd; the pressure
    target counts only when the pressure is corrected."""
    mesh = _tiny_cylinder()
    fem, ctx, _ = _ctx(mesh)
    L = sfem.Learner(ctx, 16, seed=7)
    L.set_components(ncomp)
    snaps = _record_random_trajectory(ctx, L, mesh, 7, seed=2)
    assert L.info()[:2] == (7, 4)
    w = _random_weights(13, 0.3)
    L.set_weights(w)
    rng = np.random.default_rng(4)
    n = 150  # three workgroups, the last one partly filled
    windows = rng.integers(2, 6, size=n)
    nodes = rng.integers(0, mesh.nNode, size=n)
    nodes[:5] = np.asarray(mesh.dbc_node[:5])  # prescribed components are masked out
    g, loss = L.gradient(windows, nodes)
    lref, gref = _torch_loss_grad(w, snaps, windows, nodes, _neighbours(mesh), _free_mask(mesh), ncomp)
    assert loss == pytest.approx(lref, rel=1e-4)
    assert np.max(np.abs(g - gref)) <= 1e-4 * np.max(np.abs(gref))
    assert np.all(g[SL["tau"][0]:] == 0)
    for k, (o, s) in SL.items():  # each block's gradient separately, so no block can hide
        if k == "tau":
            continue
        blk = slice(o, o + int(np.prod(s)))
        assert np.linalg.norm(g[blk] - gref[blk]) <= 1e-4 * np.linalg.norm(gref[blk]) + 1e-12, k


def _mix64(z):
    M = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_adam_step_on_the_hashed_minibatch():
    """One training step samples its minibatch with the documented splitmix64 rule and
    moves each weight by lr * g / (|g| + eps) (the first bias-corrected Adam step)."""
    mesh = _tiny_cylinder()
    fem, ctx, _ = _ctx(mesh)
    seed = 12345
    L = sfem.Learner(ctx, 16, seed=seed)
    _record_random_trajectory(ctx, L, mesh, 6, seed=3)
    w0 = _random_weights(17, 0.3)
    L.set_weights(w0)
    batch, lr = 200, 1e-3
    wins = np.array([2, 3, 4])  # windows of a 6-state trajectory
    sel = [_mix64(seed ^ _mix64(0 * (1 << 32) + k)) for k in range(batch)]
    windows = np.array([wins[(h >> 32) % len(wins)] for h in sel])
    nodes = np.array([(h & 0xFFFFFFFF) % mesh.nNode for h in sel])
    g, lref = L.gradient(windows, nodes)
    loss = L.train(1, batch, lr)
    w1 = L.get_weights()
    assert loss[0] == pytest.approx(lref, rel=1e-5)
    step = w0.astype(np.float64) - w1
    expect = lr * g / (np.abs(g) + 1e-8)
    assert np.max(np.abs(step - expect)) <= 2e-6, np.max(np.abs(step - expect))
    assert np.array_equal(w1[SL["tau"][0]:], TAU)  # set_weights fixed the output scale
    assert L.info()[2] == 1


@pytest.mark.parametrize("ncomp", [2, 3])
def test_learned_start_never_changes_the_answer(ncomp):
    """Train briefly on a short trajectory, then rerun it with the learned predictor
    (correcting u, v or u, v, p):
    each step starts from the smaller-residual candidate, and every converged state
    agrees with the extrapolated run's within 1e-8 of the state's largest entry.  Both
    runs iterate Newton to ||F|| < 5e-12 (Krylov rtol 1e-10 per solve; the oracle's
    direct-solve Newton iteration levels off at 1e-13 to 3e-13 on this mesh), so the
    difference left is that of two solutions of the same discrete system at that
    residual level."""
    mesh = _tiny_cylinder()
    nsteps, tol = 14, 5e-12
    fem, ctx, opts = _ctx(mesh)
    ctx.set_predictor(sfem.PRED_EXTRAPOLATE)
    L = sfem.Learner(ctx, nsteps + 1, seed=21)
    L.set_components(ncomp)
    L.record(ctx, new_trajectory=True)
    ref, its_ref = [], []
    t = 0.1
    for s in range(nsteps):
        st = ctx.step(min(t, 1.0), opts, tol, 10, allow_notconv=False)
        its_ref.append(st.krylov_iters)
        L.record(ctx)
        ref.append(ctx.get_solution())
        t += 0.1
    loss = L.train(300, 1024, 2e-3)
    assert np.all(np.isfinite(loss)) and loss[-20:].mean() < loss[:20].mean()
    tau = L.get_weights()[SL["tau"][0]:]
    assert np.all(np.isfinite(tau)) and np.all(tau > 0)
    fem2, ctx2, _ = _ctx(mesh)
    ctx2.set_learner(L)
    ctx2.set_predictor(sfem.PRED_LEARNED)
    t = 0.1
    picked = []
    for s in range(nsteps):
        st = ctx2.step(min(t, 1.0), opts, tol, 10, allow_notconv=False)
        cnt, nrm = ctx2.predictor_stats()
        picked.append(nrm)
        assert st.rhs_norm0 == pytest.approx(min(nrm), rel=1e-6)
        sol = ctx2.get_solution()
        scale = np.max(np.abs(ref[s]))
        print(f"[learned] step {s}: start {['linear', 'quadratic', 'learned'][int(np.argmin(nrm))]}, "
              f"newton {st.newton_iters}, |F| {st.rhs_norm:.2e}, max |diff| / max |ref| "
              f"{np.max(np.abs(sol - ref[s])) / scale:.2e}", flush=True)
        assert np.max(np.abs(sol - ref[s])) <= 1e-8 * scale, (s, np.max(np.abs(sol - ref[s])) / scale)
        t += 0.1
    cnt, _ = ctx2.predictor_stats()
    assert cnt.sum() == nsteps


@pytest.mark.parametrize("ncomp", [2, 3])
def test_output_scale_from_hashed_samples(ncomp):
    """Before the first Adam step tau_c is set to the RMS of the scaled targets over the
    65,536 samples of minibatch tag 2^32 - 1 (free components that are corrected)."""
    mesh = _tiny_cylinder()
    fem, ctx, _ = _ctx(mesh)
    seed = 99
    L = sfem.Learner(ctx, 16, seed=seed)
    L.set_components(ncomp)
    snaps = _record_random_trajectory(ctx, L, mesh, 6, seed=5)
    L.train(1, 64, 1e-3)
    tau = L.get_weights()[SL["tau"][0]:].astype(np.float64)
    wins = [2, 3, 4]
    tag = 0xFFFFFFFF
    mask = _free_mask(mesh).reshape(-1, 3)
    sq, cnt = np.zeros(3), np.zeros(3)
    per_win = {}
    nbrs = _neighbours(mesh)
    for t in wins:
        s0, s1, s2, s3 = (snaps[t + d].astype(np.float64).reshape(-1, 3) for d in (0, -1, -2, 1))
        dd = s0 - s1
        isig = (np.float32(1.0) / np.sqrt(np.mean(dd ** 2, axis=0)).astype(np.float32)).astype(np.float64)
        _, _, _, lam = _features(s0, s1, s2, nbrs, isig)
        tt = ((s3 - s0 - 2 * dd + (s1 - s2)) * isig).astype(np.float32)
        per_win[t] = (tt / lam[:, None].astype(np.float32)).astype(np.float64)
    for k in range(65536):
        h = _mix64(seed ^ _mix64(tag * (1 << 32) + k))
        t, i = wins[(h >> 32) % len(wins)], (h & 0xFFFFFFFF) % mesh.nNode
        sq += mask[i] * per_win[t][i] ** 2
        cnt += mask[i]
    ref = np.sqrt(sq[:ncomp] / cnt[:ncomp])
    assert np.all(ref > 0)
    assert np.allclose(tau[:ncomp], ref, rtol=1e-5), (tau, ref)
    if ncomp == 2:
        assert tau[2] == 1.0  # the pressure is not corrected: no samples, unit scale
