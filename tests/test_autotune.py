"""The sort autotuner's cache (host/autotune.cpp) and its digit plans.

CPU: entries stored in one process are read back by another through the
cache file (RWH_TUNE_CACHE), per size bucket, with batches under their own
entries; the file of the earlier format (a bare tile shape per bucket) reads
as digit plan 0; entries that name no compiled shape or plan are ignored.
GPU: every tile shape with every digit plan sorts exactly, and the tuner's
pick is a valid configuration."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def in_fresh_process(code, cache):
    env = dict(os.environ, RWH_TUNE_CACHE=str(cache))
    prog = ("import sys; sys.path.insert(0, %r)\n"
            "import __graft_entry__ as ge\nrw = ge.load_package()\n" % ROOT) + code
    r = subprocess.run([sys.executable, "-c", prog], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip()


def test_plans_cover_the_morton_bits(rw):
    plans = rw.sort_plans()
    assert plans[0] == (8, 8, 8, 6)
    assert len(set(plans)) == len(plans) >= 4
    for w in plans:
        assert sum(w) == 30 and all(1 <= x <= 8 for x in w) and w[0] >= 6


def test_cache_round_trip(rw, tmp_path):
    cache = tmp_path / "tune.json"
    in_fresh_process("rw.tune_store(1 << 24, 5, 1)\nrw.tune_store(1 << 20, 2, 3, chunk=True)\n"
                     "rw.tune_store(5000, 0, 4)\n", cache)
    data = json.load(open(cache))
    assert data == {"sort": {"24": [5, 1], "chunk/20": [2, 3], "13": [0, 4]}}
    out = in_fresh_process(
        "print(rw.tune_lookup(1 << 24), rw.tune_lookup(15000000), rw.tune_lookup(1 << 20, chunk=True),"
        " rw.tune_lookup(1 << 20), rw.tune_lookup(8000))", cache)
    assert out == "(5, 1) (5, 1) (2, 3) None (0, 4)"


def test_cache_reads_the_earlier_format_and_skips_bad_entries(rw, tmp_path):
    cache = tmp_path / "tune.json"
    cache.write_text('{"sort": {"24": 5, "22": [99, 0], "21": [1, 77], "20": [3, 2]}}\n')
    out = in_fresh_process("print(rw.tune_lookup(1 << 24), rw.tune_lookup(1 << 22), rw.tune_lookup(1 << 21),"
                           " rw.tune_lookup(1 << 20))", cache)
    assert out == "(5, 0) None None (3, 2)"


@pytest.mark.gpu
@pytest.mark.parametrize("n", [4097, 100000, 1 << 20])
def test_sort_every_shape_and_plan(gpu, n):
    rw = gpu
    L = rw.lib()
    rng = np.random.default_rng(n + 3)
    k = rng.integers(0, 1 << 30, size=n, dtype=np.uint64).astype(np.uint32)
    want = np.sort(k)
    d_in = rw.DeviceArray.from_numpy(k)
    d_out = rw.DeviceArray(n, np.uint32)
    ws = rw.Workspace(n)
    for v in range(len(rw.sort_variants())):
        for p in range(len(rw.sort_plans())):
            print(f"case sort n={n} shape={v} plan={p}", flush=True)
            ws.set_sort_variant(v)
            ws.set_sort_plan(p)
            rw.check(L.rwh_sort_u32(ws.ptr, d_in.ptr, d_out.ptr, n, 30, None))
            assert np.array_equal(d_out.numpy(), want), (v, p)


@pytest.mark.gpu
def test_tuner_picks_a_valid_configuration(gpu, tmp_path):
    rw = gpu
    out = in_fresh_process("print(*rw.best_sort_config(1 << 20), *rw.best_sort_config(1 << 20, chunk=True))",
                           tmp_path / "tune.json")
    v, p, vc, pc = (int(x) for x in out.split())
    shapes = rw.sort_variants()
    assert 0 <= v < len(shapes) and 0 <= p < len(rw.sort_plans())
    assert shapes[vc][0] * shapes[vc][1] <= 8192 and 0 <= pc < len(rw.sort_plans())
    data = json.load(open(tmp_path / "tune.json"))
    assert data["sort"]["20"] == [v, p] and data["sort"]["chunk/20"] == [vc, pc]
