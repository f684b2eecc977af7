"""The autotuner (cilreg_autotune, opts.autotune).

It times candidate tilings and temporal depths and keeps the fastest; since
the results do not depend on either, a tuned solve must equal the default
solve bit for bit, for single solves and for batches, and the choice must be
one of the candidates and be cached (a second query returns it at once).
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["rof", "fgp", "pd", "tgv"])
def test_autotune_choice_is_a_candidate_and_cached(method):
    from ccpi.filters import device as D
    shape = (24, 40, 72)
    c = D.autotune(method, shape)
    assert c["tile_y"] in (4, 8)
    assert c["zchunk"] in (0, 8, 16)
    assert c["depth"] in ((1, 2) if method != "rof" else (1,))
    assert c["ms_per_iter"] > 0
    t0 = time.perf_counter()
    again = D.autotune(method, shape)
    assert again == c
    assert time.perf_counter() - t0 < 0.05
    two_d = D.autotune(method, (100, 77))
    assert two_d["tile_y"] == 4 and two_d["depth"] == (2 if method == "fgp" else 1)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_tuned_solves_equal_default_solves(precision):
    from ccpi.filters import batch as B
    from ccpi.filters import device as D
    x = B.phantom((20, 52, 90), 5, 0.05, 77)
    pairs = [
        (D.fgp(x, 0.03, 13, precision=precision), D.fgp(x, 0.03, 13, precision=precision, autotune=True)),
        (D.tgv(x, 0.03, 1.0, 2.0, 11, precision=precision),
         D.tgv(x, 0.03, 1.0, 2.0, 11, precision=precision, autotune=True)),
        (D.pd(x, 0.03, 9, precision=precision), D.pd(x, 0.03, 9, precision=precision, autotune=True)),
        (D.rof(x, 0.03, 15, 0.002), D.rof(x, 0.03, 15, 0.002, autotune=True)),
    ]
    for (a, ia), (b, ib) in pairs:
        assert np.array_equal(ia, ib)
        assert bool((a == b).all())


def test_tuned_batch_equals_single_solves():
    from ccpi.filters import batch as B
    from ccpi.filters import device as D
    shapes = [(20, 37, 70), (9, 64, 65), (33, 17, 130), (11, 40, 72), (100, 77)]
    methods = ["fgp", "rof", "pd", "tgv", "fgp"]
    xs = [B.phantom(s, 5, 0.05, 400 + i) for i, s in enumerate(shapes)]
    specs = []
    for m, x in zip(methods, xs):
        sp = {"method": m, "input": x, "iters": 17, "lambda": 0.03, "tau": 0.002, "L": 12.0}
        specs.append(sp)
    got = B.batch_run(specs, autotune=True)
    for (m, x), (y, info) in zip(zip(methods, xs), got):
        if m == "fgp":
            ref = D.fgp(x, 0.03, 17)
        elif m == "rof":
            ref = D.rof(x, 0.03, 17, 0.002)
        elif m == "pd":
            ref = D.pd(x, 0.03, 17, L=12.0)
        else:
            ref = D.tgv(x, 0.03, 1.0, 2.0, 17)
        assert np.array_equal(info, ref[1]), m
        assert bool((y == ref[0]).all()), m


def test_tuning_cache_file_round_trip(tmp_path):
    """Choices saved to a file and loaded into a cleared cache are used
    without timing anything; the entry names this device."""
    from ccpi.filters import device as D
    shape = (28, 44, 80)
    first = D.autotune("fgp", shape, precision="bf16")
    path = tmp_path / "tune.txt"
    assert D.tune_save(path) >= 1
    rows = [l.split() for l in open(path) if not l.startswith("#")]
    assert all(r[0].startswith("gfx") and ":" in r[0] for r in rows)
    D.tune_clear()
    assert D.tune_entries() == 0
    assert D.tune_load(path) == len(rows)
    t0 = time.perf_counter()
    again = D.autotune("fgp", shape, precision="bf16")
    assert time.perf_counter() - t0 < 0.05
    assert again == first


def test_tuning_cache_environment_variable(tmp_path):
    """CILREG_TUNE_CACHE: a fresh process records its choices there, and a
    second process reuses them without timing."""
    import json
    import os
    import subprocess
    import sys
    path = tmp_path / "env_tune.txt"
    code = ("import json, sys, time; sys.path.insert(0, %r); from ccpi.filters import device as D; "
            "t = time.perf_counter(); c = D.autotune('pd', (30, 40, 70)); "
            "print(json.dumps([c, time.perf_counter() - t]))")
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "ccpi-regularisation-toolkit_amd")
    env = dict(os.environ, CILREG_TUNE_CACHE=str(path))
    runs = []
    for _ in range(2):
        r = subprocess.run([sys.executable, "-c", code % pkg], env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        runs.append(json.loads(r.stdout.strip().splitlines()[-1]))
    assert path.exists() and len([l for l in open(path) if not l.startswith("#")]) >= 1
    assert runs[1][0] == runs[0][0]
    assert runs[1][1] < 0.05
