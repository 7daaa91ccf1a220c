"""The plan autotuner's host logic (no GPU): which candidates it times, that it keeps the fastest,
and that the persisted cache is keyed by device, library build and plan key and is reused."""
import json

import ffs
from ffs import tune


def params(n=64, **kw):
    kw.setdefault("nu", 1e-3)
    return ffs.Params(nx=n, ny=n, nz=n, **kw)


BASE = {"x_pair": 1, "x_waves": 2, "x_split": 1, "fusion": 1}


def test_candidates_cover_the_fields_that_matter():
    none = {"x_pair": False, "x_waves": False, "x_split": False, "fusion": False}
    assert tune.candidates(none, BASE) == [BASE]
    allm = {"x_pair": True, "x_waves": True, "x_split": True, "fusion": True}
    c = tune.candidates(allm, BASE)
    # split: 2 pair x 2 waves x 2 fusion; joint general pass: pair and waves do not apply
    assert len(c) == 8 + 2
    assert len({json.dumps(x, sort_keys=True) for x in c}) == len(c)
    assert all(x["x_pair"] == BASE["x_pair"] and x["x_waves"] == BASE["x_waves"] for x in c if x["x_split"] == 0)
    fp32_plain = {"x_pair": False, "x_waves": True, "x_split": False, "fusion": True}
    c = tune.candidates(fp32_plain, {**BASE, "x_pair": 0, "fusion": 0})
    assert sorted((x["x_waves"], x["fusion"]) for x in c) == [(1, 0), (1, 1), (2, 0), (2, 1)]
    assert all(x["x_pair"] == 0 for x in c)
    # z slabs: each combination with the run's default slab size and without slabs
    c = tune.candidates({**fp32_plain, "z_slab": 32}, {**BASE, "x_pair": 0, "fusion": 0, "z_slab": 32})
    assert sorted((x["x_waves"], x["fusion"], x["z_slab"]) for x in c) == sorted(
        (w, f, z) for w in (1, 2) for f in (0, 1) for z in (0, 16, 32))
    c = tune.candidates({**fp32_plain, "z_slab": 0}, {**BASE, "z_slab": 0})
    assert len(c) == 4 and all(x["z_slab"] == 0 for x in c)


def test_candidates_with_the_looping_x_pass():
    """x_pair option 2: candidates 2, 1 and 0; the looping form (2) once per fusion choice, whatever x_waves."""
    opts = {"x_pair": 2, "x_waves": True, "x_split": False, "fusion": True}
    c = tune.candidates(opts, {**BASE, "x_pair": 2})
    assert len({json.dumps(x, sort_keys=True) for x in c}) == len(c)
    assert sorted((x["x_pair"], x["x_waves"]) for x in c if x["fusion"] == 1) == [(0, 1), (0, 2), (1, 1), (1, 2),
                                                                                 (2, 2)]
    assert sum(x["x_pair"] == 2 for x in c) == 2


def test_plan_key_separates_what_changes_the_kernels():
    k = tune.plan_key(params())
    assert k != tune.plan_key(params(precision="fp32"))
    assert k != tune.plan_key(params(les_model="Smagorinsky"))
    assert k != tune.plan_key(params(n=128))
    assert k == tune.plan_key(params(nu=5e-3))  # the viscosity does not change the kernels


def test_tune_keeps_the_fastest_and_caches_it(tmp_path):
    path = str(tmp_path / "plans.json")
    opts = {"x_pair": True, "x_waves": True, "x_split": False, "fusion": False}
    cost = {(1, 2): 1.0, (0, 2): 0.7, (1, 1): 0.9, (0, 1): 0.8}
    calls = []

    def measure(plan, steps):
        calls.append((plan["x_pair"], plan["x_waves"], steps))
        return cost[(plan["x_pair"], plan["x_waves"])]

    best = tune.tune(params(), cache=path, device="dev A", lib="lib1", measure=measure, options=opts, base=BASE)
    assert (best["x_pair"], best["x_waves"]) == (0, 2)
    n_timed = len(calls)
    assert n_timed == 1 + 2 * 4  # a probe, then two interleaved rounds over four candidates
    data = json.load(open(path))
    entry = data["dev A"]["lib1"][tune.plan_key(params())]
    assert entry["plan"] == best and len(entry["ms_per_step"]) == 4
    # the cache answers without timing; another device or library build tunes afresh
    again = tune.tune(params(), cache=path, device="dev A", lib="lib1", measure=measure, options=opts, base=BASE)
    assert again == best and len(calls) == n_timed
    tune.tune(params(), cache=path, device="dev B", lib="lib1", measure=measure, options=opts, base=BASE)
    tune.tune(params(), cache=path, device="dev A", lib="lib2", measure=measure, options=opts, base=BASE)
    assert len(calls) == 3 * n_timed
    data = json.load(open(path))
    assert set(data) == {"dev A", "dev B"} and set(data["dev A"]) == {"lib1", "lib2"}


def test_a_single_candidate_is_not_timed(tmp_path):
    path = str(tmp_path / "plans.json")
    none = {"x_pair": False, "x_waves": False, "x_split": False, "fusion": False}

    def measure(plan, steps):
        raise AssertionError("nothing to time")

    assert tune.tune(params(), cache=path, device="d", lib="l", measure=measure, options=none, base=BASE) == BASE
