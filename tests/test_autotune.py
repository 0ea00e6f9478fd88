"""Autotuner: launch shapes change speed, never routes."""
import pytest

from instantgr_amd import autotune


def test_cache_roundtrip(tmp_path):
    p = str(tmp_path / "tune.json")
    assert autotune.load(p, "k") is None
    res = {"tuning": {"sched_group": 8, "sched_chunk": 16, "graph": 1, "dag_order": 1, "dag_waves": 2,
                      "track_k": 4, "dag_edges": 1, "dag_seq": 8},
           "trials": []}
    autotune.save(p, "k", res)
    autotune.save(p, "other", {"tuning": {"graph": 2}, "trials": []})
    assert autotune.load(p, "k") == res["tuning"]
    assert autotune.load(p, "other") == {"graph": 2}


def test_candidates_cover_fields():
    from instantgr_amd import gpu
    assert set(autotune.CANDIDATES) | autotune.PLAIN_ONLY == set(gpu.TUNING_FIELDS)
    assert set(autotune.STAGE) | autotune.PLAIN_ONLY == set(gpu.TUNING_FIELDS)


@pytest.mark.gpu
def test_tune_small_design_keeps_routes():
    from instantgr_amd import gpu
    from instantgr_amd.host import Design
    d = Design.preset(0, nets=3000).build(8)
    res = autotune.tune(d, rrr=1, threads=8,
                        candidates={"sched_group": [4, 16], "sched_chunk": [4, 32], "graph": [1, 2],
                                    "dag_order": [1, 2], "dag_waves": [2, 4], "track_k": [4, 8],
                                    "dag_edges": [1, 2], "dag_seq": [2, 1 << 30], "sched_kind": [1, 2],
                                    "dag_grid": [16, 2048], "lshape_groups": [1, 4]})
    t = res["tuning"]
    assert t["sched_group"] in (4, 16) and t["sched_chunk"] in (4, 32)
    assert t["graph"] in (1, 2) and t["dag_order"] in (1, 2) and t["dag_waves"] in (2, 4)
    assert t["track_k"] in (4, 8) and t["dag_grid"] in (16, 2048) and t["lshape_groups"] in (1, 4)
    assert all(tr["routes_match"] for tr in res["trials"])
    # the chosen shape, applied, routes as the built-in one does
    r0 = gpu.Router.from_design(d)
    r0.run(d, form=gpu.FORM_BATCHED, rrr=1, threads=8)
    r1 = gpu.Router.from_design(d)
    autotune.apply(r1, t)
    assert r1.tuning()["graph"] == t["graph"]
    r1.run(d, form=gpu.FORM_BATCHED, rrr=1, threads=8)
    assert gpu.routes_equal(r0.routes_flat(), r1.routes_flat())
    r0.close()
    r1.close()


@pytest.mark.gpu
def test_tuning_rejects_bad_values():
    from instantgr_amd import gpu
    from instantgr_amd.host import Design
    d = Design.preset(0, nets=200).build(4)
    r = gpu.Router.from_design(d)
    with pytest.raises(RuntimeError):
        r.set_tuning(sched_group=5)
    with pytest.raises(RuntimeError):
        r.set_tuning(sched_chunk=1000)
    with pytest.raises(KeyError):
        r.set_tuning(tile=3)
    r.close()
