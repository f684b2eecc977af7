"""Grouped multi-volume execution on the GPU: the batch runner and the
continuous-batching server.

The property under test is the one the engine is built on: a volume's result
and infovector are bit-identical whether it is solved alone (the device API,
itself pinned bit-exactly to the oracle by test_gpu_parity.py) or inside a
batch / the server, in any mix of methods, dimensionalities, sizes, iteration
counts, tolerances and admission times.  Single solves use zchunk=32, the
batch's plane count, so tolerance-stopped jobs see the same per-block partial
sums.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mix():
    # (method, shape, params): 2D and 3D, uneven extents, all four methods,
    # both TV types, nonneg, fixed counts and tolerance stops
    return [
        ("fgp", (20, 37, 70), dict(lam=0.03, iters=23, tol=0.0, method=0, nonneg=1)),
        ("rof", (9, 64, 65), dict(lam=0.04, iters=41, tau=0.002, tol=0.0)),
        ("pd", (33, 17, 130), dict(lam=0.02, iters=17, tol=0.0, L=12.0, method=1, nonneg=0)),
        ("tgv", (11, 40, 72), dict(lam=0.05, iters=15, alpha1=1.0, alpha0=2.0, L2=12.0, tol=0.0)),
        ("fgp", (100, 77), dict(lam=0.05, iters=60, tol=2e-3, method=1, nonneg=0)),
        ("rof", (256, 130), dict(lam=0.02, iters=90, tau=0.001, tol=1e-3)),
        ("pd", (64, 64), dict(lam=0.03, iters=44, tol=1e-3, L=8.0, method=0, nonneg=1)),
        ("tgv", (96, 50), dict(lam=0.04, iters=31, alpha1=1.0, alpha0=2.0, L2=12.0, tol=0.0)),
        ("fgp", (16, 70, 66), dict(lam=0.02, iters=35, tol=1e-3, method=0, nonneg=0)),
        ("tgv", (7, 33, 129), dict(lam=0.03, iters=12, alpha1=1.0, alpha0=2.0, L2=12.0, tol=1e-2)),
        ("rof", (5, 5), dict(lam=0.04, iters=9, tau=0.002, tol=0.0)),
    ]


def _spec(method, x, p):
    s = {"method": method, "input": x, "iters": p["iters"], "tol": p["tol"], "lambda": p["lam"]}
    if method == "rof":
        s["tau"] = p["tau"]
    if method in ("fgp", "pd"):
        s["method_tv"] = p["method"]
        s["nonneg"] = p["nonneg"]
    if method == "pd":
        s["L"] = p["L"]
    if method == "tgv":
        s.update(alpha1=p["alpha1"], alpha0=p["alpha0"], L2=p["L2"])
    return s


def _single(method, x, p, **kw):
    from ccpi.filters import device as D
    kw.setdefault("zchunk", 32)
    if method == "rof":
        return D.rof(x, p["lam"], p["iters"], p["tau"], p["tol"], **kw)
    if method == "fgp":
        return D.fgp(x, p["lam"], p["iters"], p["tol"], p["method"], p["nonneg"], **kw)
    if method == "pd":
        return D.pd(x, p["lam"], p["iters"], p["tol"], p["L"], p["method"], p["nonneg"], **kw)
    return D.tgv(x, p["lam"], p["alpha1"], p["alpha0"], p["iters"], p["L2"], p["tol"], **kw)


def _inputs(mix, seed0=100):
    from ccpi.filters import batch as B
    return [B.phantom(shape, n_shapes=6, sigma=0.05, seed=seed0 + i) for i, (_, shape, _) in enumerate(mix)]


def _assert_same(method, shape, got, ref):
    (a, ia), (b, ib) = got, ref
    tag = f"{method} {shape}"
    assert np.array_equal(ia, ib), f"{tag}: infovector {ia} != {ib}"
    assert a.shape == b.shape
    assert bool((a == b).all()), f"{tag}: {int((a != b).sum())} differing elements"


PERSISTENT_OFF = ("the persistent work-queue batch mode outlived its test limit on the GPU twice (cause not "
                  "established); the library refuses it and it stays off the GPU")


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("mode", ["graph", "stream",
                                  pytest.param("persistent", marks=pytest.mark.skip(reason=PERSISTENT_OFF))])
def test_batch_matches_single_solves(precision, mode):
    """In persistent mode the groups whose jobs all run fixed counts (3D ROF,
    3D PD, 2D TGV here) run as one work-queue kernel each and the others keep
    per-level launches, in the same batch."""
    from ccpi.filters import batch as B
    mix = _mix()
    xs = _inputs(mix)
    specs = [_spec(m, x, p) for (m, _, p), x in zip(mix, xs)]
    print(f"batch of {len(specs)} jobs, {precision}, {mode} launches", flush=True)
    got = B.batch_run(specs, precision=precision, stream_launches=mode == "stream", persistent=mode == "persistent")
    for (m, shape, p), x, g in zip(mix, xs, got):
        _assert_same(m, shape, g, _single(m, x, p, precision=precision))
    # at least one tolerance job stopped early, so the stop path is covered
    stopped = [int(g[1][0]) < p["iters"] for (m, _, p), g in zip(mix, got) if p["tol"] > 0]
    assert any(stopped), [g[1] for g in got]


@pytest.mark.skip(reason=PERSISTENT_OFF)
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_persistent_batch_fixed_counts(precision):
    """Every group persistent (all jobs fixed counts): two-iteration passes
    with odd tails (3D FGP, PD, TGV), one-iteration groups (ROF, 2D), jobs of
    different lengths sharing a queue; each result bit-identical to its single
    solve, repeated to show the queues reset."""
    from ccpi.filters import batch as B
    mix = [(m, shape, dict(p, tol=0.0)) for m, shape, p in _mix()]
    xs = _inputs(mix, seed0=300)
    specs = [_spec(m, x, p) for (m, _, p), x in zip(mix, xs)]
    for _ in range(2):
        got = B.batch_run(specs, precision=precision, persistent=True)
        for (m, shape, p), x, g in zip(mix, xs, got):
            _assert_same(m, shape, g, _single(m, x, p, precision=precision))


def test_batch_generated_inputs_match_phantom_api():
    from ccpi.filters import batch as B
    mix = _mix()[:6]
    specs = []
    refs = []
    for i, (m, shape, p) in enumerate(mix):
        s = _spec(m, None, p)
        del s["input"]
        s.update(shape=shape, phantom=(5, 0.04, 7000 + i))
        specs.append(s)
        refs.append(_single(m, B.phantom(shape, 5, 0.04, 7000 + i), p))
    got = B.batch_run(specs)
    for (m, shape, _), g, r in zip(mix, got, refs):
        _assert_same(m, shape, g, r)


def test_batch_zero_iterations_and_empty():
    import torch
    from ccpi.filters import batch as B
    assert B.batch_run([]) == []
    x = B.phantom((6, 20, 70), 4, 0.05, 3)
    sentinel = torch.full_like(x, 7.0)
    got = B.batch_run([{"method": "rof", "input": x, "iters": 0, "lambda": 0.02, "tau": 0.001},
                       {"method": "fgp", "input": x, "iters": 0, "lambda": 0.02}],
                      outs=[torch.zeros_like(x), sentinel])
    assert torch.equal(got[0][0], x)            # ROF: Output = Input
    assert bool((got[1][0] == 7.0).all())       # FGP: the output is left as it was
    assert got[0][1][0] == 0 and got[1][1][0] == 0


def test_server_results_independent_of_the_mix():
    from ccpi.filters import batch as B
    mix = _mix()
    xs = _inputs(mix, seed0=300)
    refs = [_single(m, x, p) for (m, _, p), x in zip(mix, xs)]
    with B.Server(0, hbm_budget=8e9) as srv:
        tickets = {}
        # staggered admission: a few jobs, some levels, then the rest
        for i in range(4):
            tickets[i] = srv.submit(_spec(mix[i][0], xs[i], mix[i][2]))
        srv.step(5)
        for i in range(4, len(mix)):
            tickets[i] = srv.submit(_spec(mix[i][0], xs[i], mix[i][2]))
            srv.step(3)
        results = {}
        for _ in range(1000):
            for i, t in tickets.items():
                if i not in results:
                    r = srv.poll(t)
                    if r is not None:
                        results[i] = r
            if len(results) == len(mix):
                break
            srv.step(4)
        assert len(results) == len(mix)
    for i, (m, shape, _) in enumerate(mix):
        _assert_same(m, shape, results[i], refs[i])


def test_server_budget_queues_and_rejects():
    from ccpi.filters import batch as B
    mix = _mix()[:4]
    xs = _inputs(mix, seed0=500)
    refs = [_single(m, x, p) for (m, _, p), x in zip(mix, xs)]
    # room for two of these jobs at a time (each needs 0.3-3.3 MB)
    with B.Server(0, hbm_budget=4e6) as srv:
        ts = [srv.submit(_spec(m, x, p)) for (m, _, p), x in zip(mix, xs)]
        big = srv.submit({"method": "fgp", "shape": (64, 256, 256), "phantom": (4, 0.05, 1), "iters": 5,
                          "lambda": 0.02})
        active, pending = srv.step(1)
        assert active >= 1 and pending >= 1, (active, pending)
        done = {}
        for _ in range(2000):
            for i, t in enumerate(ts):
                if i not in done:
                    r = srv.poll(t)
                    if r is not None:
                        done[i] = r
            if len(done) == len(ts):
                break
            srv.step(8)
        assert len(done) == len(ts)
        with pytest.raises(RuntimeError):
            srv.poll(big)
    for i, (m, shape, _) in enumerate(mix):
        _assert_same(m, shape, done[i], refs[i])


def _phantom_numpy(shape, n_shapes, sigma, seed):
    """Restatement of the device phantom generator (splitmix64 streams, float32
    geometry, Box-Muller noise) for a small volume."""
    M = (1 << 64) - 1

    def mix(s):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return s, z ^ (z >> 31)

    def unit(v):
        return (v >> 11) * (1.0 / 9007199254740992.0)

    nz, ny, nx = shape if len(shape) == 3 else (1,) + tuple(shape)
    is2d = len(shape) == 2
    st = (seed * 0x2545F4914F6CDD1D + 1) & M
    shapes = []
    for _ in range(n_shapes):
        c, r = [], []
        for _ in range(3):
            st, v = mix(st)
            c.append(np.float32(unit(v)))
        for _ in range(3):
            st, v = mix(st)
            r.append(np.float32(0.05) + np.float32(0.25) * np.float32(unit(v)))
        st, v = mix(st)
        val = np.float32(unit(v))
        rect = 0
        if is2d:
            st, v = mix(st)
            rect = int(unit(v) < 0.5)
        shapes.append((c, r, val, rect))
    out = np.zeros(nz * ny * nx, dtype=np.float32)
    for i in range(out.size):
        x, y, z = i % nx, (i // nx) % ny, i // (nx * ny)
        p = [(np.float32(x) + np.float32(0.5)) / np.float32(nx), (np.float32(y) + np.float32(0.5)) / np.float32(ny),
             (np.float32(z) + np.float32(0.5)) / np.float32(nz)]
        v = np.float32(0)
        for c, r, val, rect in shapes:
            if rect:
                inside = abs(p[0] - c[0]) <= r[0] and abs(p[1] - c[1]) <= r[1] and (nz == 1 or abs(p[2] - c[2]) <= r[2])
            else:
                q = np.float32(0)
                for d in range(2 if nz == 1 else 3):
                    t = (p[d] - c[d]) / r[d]
                    q = np.float32(q + t * t)
                inside = q <= 1
            if inside and val > v:
                v = val
        s = (seed ^ ((0xD1B54A32D192ED03 * (i + 1)) & M)) & M
        s, a = mix(s)
        s, b = mix(s)
        u1 = max(unit(a), 1e-300)
        g = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * unit(b))
        out[i] = v + np.float32(sigma) * np.float32(g)
    return out.reshape(shape)


@pytest.mark.parametrize("shape", [(5, 6, 7), (9, 11)], ids=["3d", "2d"])
def test_phantom_matches_restatement(shape):
    from ccpi.filters import batch as B
    got = B.phantom(shape, 4, 0.1, 1234).cpu().numpy()
    ref = _phantom_numpy(shape, 4, 0.1, 1234)
    # double log/cos may differ in the last bit between libraries: compare to 1e-6
    assert np.allclose(got, ref, rtol=0, atol=1e-6), np.abs(got - ref).max()
    again = B.phantom(shape, 4, 0.1, 1234).cpu().numpy()
    assert np.array_equal(got, again)


def test_persistent_mode_is_refused():
    """The disabled mode returns -1 with its reason before any launch."""
    from ccpi.filters import batch as B
    x = B.phantom((8, 20, 70), 3, 0.05, 1)
    with pytest.raises(RuntimeError, match="persistent batch launches are disabled"):
        B.batch_run([{"method": "rof", "input": x, "iters": 4, "lambda": 0.02, "tau": 0.001}], persistent=True)
