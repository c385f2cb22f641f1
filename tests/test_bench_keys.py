"""bench.py names a variant's kernel the way openpfc_amd.profile_summary
shortens the demangled name rocprofv3 prints, so the roofline's traffic
lookup finds the kernel the step actually ran.  The demangled names below are
the form rocprofv3 reports for these launches (template arguments in
declaration order, defaulted ones included)."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402
from openpfc_amd.profile_summary import short_name  # noqa: E402

PLAN8 = "pfc_amd::FftPlan<1024, 8, pfc_amd::RadixList<8, 8, 4, 4>, false>"
PLAN16 = "pfc_amd::FftPlan<1024, 16, pfc_amd::RadixList<16, 16, 4>, false>"
HALF4 = "pfc_amd::FftPlan<512, 4, pfc_amd::RadixList<4, 4, 4, 4, 2>, false>"
HALF8 = "pfc_amd::FftPlan<512, 8, pfc_amd::RadixList<8, 8, 8>, false>"
PLAN8XP = "pfc_amd::FftPlan<1024, 8, pfc_amd::RadixList<8, 8, 4, 4>, true>"
HALF8XP = "pfc_amd::FftPlan<512, 8, pfc_amd::RadixList<8, 8, 8>, true>"


def demangled(kernel, t, plan, rest):
    return f"void pfc_amd::{kernel}<{t}, {plan}, {rest}>(pfc_amd::TStepArgs<{t}>)"


@pytest.mark.parametrize("kernel,variant,mode,name", [
    ("k_ty_fwd", "r8x8x4x4/E8/G8/split/wide2/nt", 0,
     demangled("k_ty_fwd_w", "double", PLAN8, "8, true, 2, true")),
    ("k_ty_fwd", "r8x8x4x4/E8/G8/split/wide2", 0,
     demangled("k_ty_fwd_w", "double", PLAN8, "8, true, 2, false")),
    ("k_tz", "r8x8x4x4/E8/G1/full/pre/nt", 0,
     demangled("k_tz", "double", PLAN8, "1, false, 0, true, true")),
    ("k_tz", "r8x8x4x4/E8/G1/split", 2,
     demangled("k_tz", "double", PLAN8, "1, true, 2, false, false")),
    ("k_tx_fwd", "r16x16x4/E16/G8/split/wide", 0,
     demangled("k_tx_fwd_w", "double", PLAN16, "8, true, false")),
    ("k_ty_inv", "r16x16x4/E16/G8/split/wide2", 0,
     demangled("k_ty_inv_w", "double", PLAN16, "8, true, 2, false")),
    ("k_tx_inv", "r4x4x4x4x2/E4/G8/nt", 0,
     demangled("k_tx_inv", "double", HALF4, "8, true, true")),
    ("k_tx_inv", "r8x8x8/E8/G16", 0,
     demangled("k_tx_inv", "float", HALF8, "16, true, false")),
    ("k_ty_fwd", "r16x16x4/E16/G16/split", 0,
     demangled("k_ty_fwd", "float", PLAN16, "16, true, true")),
    ("k_tx_tung", "r8x8x8/E8/G4/wide", 0,
     demangled("k_tx_tung", "double", HALF8, "4, true, true")),
    ("k_tx_tung", "r8x8x8/E8/G8/wide/w4", 0,
     demangled("k_tx_tung_o", "double", HALF8, "8, true, true, 4")),
    ("k_ty_fwd", "r8x8x4x4/E8/xp/G16/split/wide2/nt", 0,
     demangled("k_ty_fwd_w", "float", PLAN8XP, "16, true, 2, true")),
    ("k_tx_inv", "r8x8x8/E8/xp/G16/nt", 0,
     demangled("k_tx_inv", "float", HALF8XP, "16, true, true")),
    ("k_tz", "r8x8x4x4/E8/xp/G1/split/nt", 0,
     demangled("k_tz", "float", PLAN8XP, "1, true, 0, false, true")),
    ("k_tz", "r8x8x4x4/E8/G1/full/mt", 0,
     demangled("k_tz_mt", "double", PLAN8, "1, false, 0, false")),
    ("k_tz", "r8x8x4x4/E8/xp/G2/split/mt", 2,
     demangled("k_tz_mt", "float", PLAN8XP, "2, true, 2, false")),
    ("k_tz", "r16x16x4/E16/G2/split/mtp", 0,
     demangled("k_tz_mt", "double", PLAN16, "2, true, 0, true")),
    ("k_tx_fuse", "r8x8x8/E8+r16x16x4/E16/G8/split/w4", 0,
     demangled("k_tx_fuse", "double", HALF8 + ", " + PLAN16, "8, true, 4")),
    ("k_tx_fuse", "r4x4x4x4x2/E4+r8x8x4x4/E8/G4/full", 0,
     demangled("k_tx_fuse", "double", HALF4 + ", " + PLAN8, "4, false, 1")),
])
def test_variant_key_matches_profile_name(kernel, variant, mode, name):
    prec = "f32" if "<float" in name else "f64"
    assert bench.variant_key(kernel, prec, 1024, variant, mode=mode) == short_name(name)


def _trace_dir(tmp_path, intervals):
    """A kernel_trace.csv with the given (start, end) nanosecond intervals."""
    import csv
    d = tmp_path / "trace"
    d.mkdir()
    with open(d / "x_kernel_trace.csv", "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for b, e in intervals:
            w.writerow(["k", b, e])
    return str(d)


def test_busy_window_after_last_gap(tmp_path):
    """Idle share after the last long gap: overlapping kernels count once,
    short gaps count as idle, the activity before the long gap is left out."""
    from openpfc_amd.profile_summary import busy_window
    ms = 1_000_000
    iv = [(0, 5 * ms),                     # before the gap (ignored)
          (300 * ms, 302 * ms), (301 * ms, 304 * ms),  # overlap: 4 ms busy
          (305 * ms, 306 * ms)]            # 1 ms idle before, 1 ms busy
    w = busy_window(_trace_dir(tmp_path, iv), 100.0)
    assert w["kernels"] == 3
    assert abs(w["span_ms"] - 6.0) < 1e-9
    assert abs(w["busy_ms"] - 5.0) < 1e-9
    assert abs(w["idle_frac"] - 1 / 6) < 1e-12

