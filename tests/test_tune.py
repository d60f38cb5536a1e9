"""The pair-kernel autotuner's host logic (candidates, choice, generated header)."""

import json
import re

import pytest

from prolix_amd import tune


def test_defaults_match_the_kernel_source():
    src = (tune.CSRC / "prx_pair_dev.h").read_text()
    for k, v in tune.DEFAULTS.items():
        m = re.search(rf"#ifndef {k}\n#define {k} (-?\d+)\n#endif", src)
        assert m, k
        assert int(m.group(1)) == v, k
        assert v in tune.KNOBS[k], k


def test_committed_header_holds_valid_settings():
    got = tune.parse_tuned()
    assert set(got) <= set(tune.KNOBS)
    for k, v in got.items():
        assert v in tune.KNOBS[k], (k, v)


def test_plan_is_coordinate_descent_around_the_base():
    base = dict(tune.DEFAULTS)
    cands = tune.plan(base)
    assert cands[0] == base
    assert len(cands) == 1 + sum(len(v) - 1 for v in tune.KNOBS.values())
    assert len({tune.name_of(c) for c in cands}) == len(cands)
    for c in cands[1:]:
        assert sum(c[k] != base[k] for k in base) == 1


def test_header_round_trip():
    s = {"PRX_PAIR_WAVES": 8, "PRX_UNROLL_BORN": 4, "PRX_UNROLL_FORCE": 2, "PRX_UNROLL_CHAIN": 4, "PRX_XCD_MAP": 0,
         "PRX_ROT_BPERM_BORN": 0, "PRX_ROT_BPERM_FORCE": 1, "PRX_ROT_BPERM_CHAIN": 1, "PRX_MIN_WAVES": 6}
    assert tune.parse_tuned(tune.header(s, "a: 1.0\nb: 2.0")) == s


@pytest.mark.parametrize("other_ms,expect_base", [(0.998, True), (0.98, False)])
def test_pick_keeps_the_base_within_noise(tmp_path, monkeypatch, other_ms, expect_base):
    monkeypatch.setattr(tune, "TUNED", tmp_path / "prx_tuned.h")
    base = tune.current()
    other = dict(base)
    other["PRX_UNROLL_FORCE"] = 8 if base["PRX_UNROLL_FORCE"] != 8 else 2
    res = tmp_path / "r.jsonl"
    res.write_text("\n".join(json.dumps({"name": tune.name_of(s), "settings": s, "ms_per_step": ms,
                                         "pass_ms": {"born_pairs": 0.4, "force_pairs": 0.6, "chain_pairs": 0.5}})
                             for s, ms in ((base, 1.0), (other, other_ms))))
    chosen, table = tune.pick(str(res))
    assert chosen == (base if expect_base else other)
    assert tune.parse_tuned((tmp_path / "prx_tuned.h").read_text()) == chosen
    assert "(chosen)" in table


def test_header_yields_to_compiler_defines():
    """A candidate's -D settings must win over the committed choice: every
    define in the generated (and the committed) header is guarded."""
    text = tune.header(dict(tune.DEFAULTS))
    for k in tune.KNOBS:
        assert f"#ifndef {k}\n#define {k} " in text
    committed = tune.TUNED.read_text()
    for k in tune.parse_tuned(committed):
        assert f"#ifndef {k}\n#define {k} " in committed, k


def test_build_makes_a_loadable_candidate(tmp_path, monkeypatch):
    """One candidate compiled by the tuner loads and exports the whole ABI."""
    import ctypes

    from tests.test_host import declared_symbols

    monkeypatch.setattr(tune, "TUNE_DIR", tmp_path)
    cand = dict(tune.DEFAULTS)
    cand["PRX_UNROLL_FORCE"] = 2
    (d,) = tune.build([cand])
    so = d / "libprolix_gb.so"
    assert so.exists() and json.loads((d / "settings.json").read_text()) == cand
    lib = ctypes.CDLL(str(so))
    for name in declared_symbols():
        assert hasattr(lib, name), name
