"""The measurement inputs bench.py reads from profiles/ stay tied to the rocprofv3 summaries they
were copied from (no GPU needed)."""
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mix():
    with open(os.path.join(ROOT, "profiles", "valu_mix.json")) as f:
        return json.load(f)["entries"]


@pytest.mark.parametrize("entry", _mix(), ids=lambda e: f"{e['family']}-{e['tile']}-{e.get('amp', False)}")
def test_valu_mix_matches_its_counter_summary(entry):
    """Every per-pair instruction count in valu_mix.json is the per-pair column of the SQ_INSTS_VALU and
    SQ_INSTS_VALU_TRANS_F32 rows of the summary it names, and that summary is about the same kernel."""
    text = open(os.path.join(ROOT, entry["source"])).read()
    assert entry["kernel"] in text
    for counter, key in (("SQ_INSTS_VALU", "valu_per_pair"), ("SQ_INSTS_VALU_TRANS_F32", "trans_per_pair")):
        m = re.search(r"^\| " + counter + r" \|[^|\n]*\| ([0-9.]+)", text, re.M)
        assert m, (entry["source"], counter)
        assert float(m.group(1)) == pytest.approx(entry[key], abs=1e-3), (entry["source"], counter)


def test_bench_reads_the_mix_of_the_plan():
    import sys

    sys.path.insert(0, ROOT)
    import bench

    m = bench.load_mix("single", 13)
    assert m["valu"] == pytest.approx(5.759) and m["trans"] == pytest.approx(0.519)
    assert bench.load_mix("single", 4) is None
    # frac counts issue slots (a transcendental = 4) at two flops per slot: at most 1 at the peak rate
    slots = m["valu"] - m["trans"] + 4 * m["trans"]
    r = {"valu": m["valu"], "trans": m["trans"], "kernel": m["kernel"], "source": m["source"]}

    class _Lib:
        @staticmethod
        def issue_rates():
            return 5.5e13, 1.8e13

    roof = bench.roofline_valu(_Lib, r, bench.PEAK_SLOTS_PER_S / slots, 23)
    assert roof["frac"] == pytest.approx(1.0)
    assert roof["algorithmic"]["frac"] > 1.0  # the reference formula's ops at that pair rate
