"""The committed rocprofv3 summaries match the launch shapes the benchmark times: for every
workload and arithmetic form with a summary, the shape recorded in the tuning cache
(tune/gfx950.txt, what bench.py's timed runs use) is the shape the summaries were taken at,
so the benchmark line's roofline figures are read from measurements of the kernel it timed."""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
PROFILES = os.path.join(ROOT, "profiles")
CACHE = os.path.join(ROOT, "shll-sve-cfd_amd", "tune", "gfx950.txt")
DEV = "gfx950:sramecc+:xnack-"

# (workload, form) -> (tuning-cache key without the device, summary prefix); precisions:
# 0 F64, 1 REF32, 2 C32, 3 C32K, 4 F64F.  Batch classes are log4 of the total cell-updates
# (tubes 2.45e12 -> 21, domains 4.29e10 -> 18), single-instance classes log2 of the cells.
CASES = {
    ("tubes", "f64f"): ("dim1 order2 prec4 batch1 class21", "k_step1d_f64f_o2_tubes4096"),
    ("tubes", "f64"): ("dim1 order2 prec0 batch1 class21", "k_step1d_f64_o2_tubes4096"),
    ("tubes", "c32k"): ("dim1 order2 prec3 batch1 class21", "k_step1d_c32k_o2_tubes4096"),
    ("tubes", "c32"): ("dim1 order2 prec2 batch1 class21", "k_step1d_c32_o2_tubes4096"),
    ("sod65536", "f64f"): ("dim1 order2 prec4 batch0 class16", "k_step1d_f64f_o2_sod65536"),
    ("sod65536", "f64"): ("dim1 order2 prec0 batch0 class16", "k_step1d_f64_o2_sod65536"),
    ("circle4096", "c32k"): ("dim2 order1 prec3 batch0 class24", "k_step2d_tb_c32k_circle4096"),
    ("circle4096", "f64"): ("dim2 order1 prec0 batch0 class24", "k_step2d_tb_f64_circle4096"),
    ("domains", "c32k"): ("dim2 order1 prec3 batch1 class18", "k_step2d_tb_c32k_domains64"),
    ("domains", "c32"): ("dim2 order1 prec2 batch1 class18", "k_step2d_tb_c32_domains64"),
    ("domains", "f64"): ("dim2 order1 prec0 batch1 class18", "k_step2d_tb_f64_domains64"),
    ("omp2d", "ref32"): ("dim2 order2mc prec1 batch0 class20", "k_step2d_ref32_o2_mc1024"),
    ("omp2d", "f64"): ("dim2 order2mc prec0 batch0 class20", "k_step2d_f64_o2_mc1024"),
    ("omp2d", "c32k"): ("dim2 order2mc prec3 batch0 class20", "k_step2d_c32k_o2_mc1024"),
}


def cache():
    out = {}
    for line in open(CACHE):
        if line.strip():
            key, val = line.strip().split("|")
            assert key not in out, "duplicate tuning-cache key " + key
            out[key] = tuple(int(x) for x in val.split())
    return out


@pytest.mark.parametrize("case", sorted(CASES))
def test_summary_shape_is_the_timed_shape(case):
    import bench
    key, prefix = CASES[case]
    c = cache()
    assert DEV + " " + key in c, "no tuning-cache line for %s" % (case,)
    shape = c[DEV + " " + key]
    for suffix, fig in (("_hbm_traffic.md", "hbm_bytes_per_cell_update"), ("_sq_counters.md", "valu_insts_per_cell_update"),
                        ("_sq_counters.md", "valu_mix_ceiling_per_simd_per_ns")):
        assert os.path.exists(os.path.join(PROFILES, prefix + suffix)), prefix + suffix
        v = bench.profile_figure(prefix + suffix, fig, shape)
        assert v is not None and v > 0, "%s: %s not taken at the timed shape %s" % (prefix + suffix, fig, shape)
    trace = open(os.path.join(PROFILES, prefix + "_kernel_trace.md")).read()
    assert re.search(r"steps/launch %d, waves/block\s+%d" % (shape[1], shape[2]), trace), prefix + " trace shape"


def test_every_summary_is_listed_in_the_index():
    index = open(os.path.join(PROFILES, "README.md")).read()
    for f in os.listdir(PROFILES):
        if f.endswith(".md") and f != "README.md":
            assert f in index, f
