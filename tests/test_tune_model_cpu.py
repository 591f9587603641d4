"""The statistics-driven tuner's cost model (owm_tune_model, host code) against
the trial times it was fitted to: on each of the recorded tile sets the
configuration the library picks from the set's leaf histogram is close to the
fastest one measured, far closer than the reference's leaf 32 with the
untuned block and key width, and the library's picks are the ones the fitting
script's least-squares model makes."""
import json
import os

import numpy as np
import pytest

from fastowm_amd import owm
from fastowm_amd.tune import fit_model as fm

HERE = os.path.dirname(os.path.abspath(__file__))
TRIALS = os.path.join(os.path.dirname(HERE), "fastowm_amd", "tune", "trials_mi355x.json")
CLASS_MAX = [1 << 16, 1 << 18, 1 << 20, 1 << 22, 1 << 24, 2 ** 64 - 1]


def cls(n):
    return next(i for i, m in enumerate(CLASS_MAX) if n <= m)


@pytest.fixture(scope="module")
def recs():
    with open(TRIALS) as f:
        return json.load(f)


def sample_of(rec):
    s = rec["sample"]
    return (rec["n"], float(rec["n"] * rec["tiles"]),
            dict(wsize=rec["wsize"], overlap=rec["overlap"], bsize=20),
            [tuple(b) for b in s["bins"]], dict(bbox=s["bbox"], max_radius=s["max_radius"]))


def library_pick(rec):
    table = owm.tune_model([sample_of(rec)])
    e_leaf = table[cls(rec["n"])]
    e_run = table[cls(rec["n"])]  # block and key width by the launch's mean tile size
    return (e_leaf["leaf_size"], (e_run["block_x"], e_run["block_y"], e_run["block_k"],
                                  e_run["summaries"]), e_run["sort_bits"])


def test_library_picks_are_near_the_fastest(recs):
    regret, base = [], []
    for rec in recs:
        t = rec["times"]
        best = min(t.values())
        regret.append(t[fm.config_key(*library_pick(rec))] / best - 1)
        base.append(t[fm.config_key(32, (8, 8, 4, 0), 32)] / best - 1)
    regret, base = np.array(regret), np.array(base)
    assert len(recs) >= 60
    assert regret.mean() < 0.03, regret.mean()
    assert regret.max() < 0.3, regret.max()
    assert np.median(regret) < 0.02
    assert base.mean() > 0.4  # the choice matters: an untuned table is far off


def test_library_agrees_with_the_fit(recs):
    coef = fm.fit(recs)
    for rec in recs:
        keys, F, _ = fm.set_matrix(rec)
        pred = F @ coef
        mine = keys.index(library_pick(rec))
        # the same pick, or one the refitted model prices within 0.1 %
        assert pred[mine] <= pred.min() * 1.001 + 1e-12, (rec["n"], rec["wsize"], keys[mine])


def test_leave_one_set_out_regret(recs):
    """The model generalises: fitted without a set, it still picks well there."""
    loo = [fm.regret(fm.fit(recs[:i] + recs[i + 1:]), recs[i]) for i in range(0, len(recs), 3)]
    assert np.mean(loo) < 0.04 and max(loo) < 0.3


def test_rejects_unusable_samples(recs):
    n, w, prm, bins, sm = sample_of(recs[0])
    with pytest.raises(owm.OWMError) as e:
        owm.tune_model([(n, n - 1.0, prm, bins, sm)])  # weight below its points
    assert e.value.name == "E_ARG"
    with pytest.raises(owm.OWMError):
        owm.tune_model([(n, w, prm, [], sm)])  # no histogram
    with pytest.raises(owm.OWMError):
        owm.tune_model([(n, w, dict(prm, overlap=1.0), bins, sm)])  # parameters derive rejects
    default = owm.tune_model([])
    assert [(e["leaf_size"], e["block_x"], e["block_y"], e["block_k"], e["summaries"], e["sort_bits"])
            for e in default] == [(128, 16, 16, 4, 0, 8)] * 3 + [(64, 16, 16, 4, 0, 16)] * 3
    assert all(e["ms"] == 0 for e in default)


def test_mixed_batch_table(recs):
    """Samples of several classes: leaf sizes per tile class, the launch choice
    in the class of the mean tile size, and a predicted time for every sampled
    class."""
    pick = [r for r in recs if r["wsize"] == 12 and r["overlap"] == 0.8 and r["cluster_frac"] == 0
            and r["density"] == 6.0]
    assert len({cls(r["n"]) for r in pick}) >= 4
    samples = [sample_of(r) for r in pick]
    table = owm.tune_model(samples)
    total = sum(s[1] for s in samples)
    tiles = sum(s[1] / s[0] for s in samples)
    for r in pick:
        assert table[cls(r["n"])]["ms"] > 0
        assert table[cls(r["n"])]["leaf_size"] in (16, 32, 64, 128)
    assert table[cls(total / tiles)]["ms"] > 0
