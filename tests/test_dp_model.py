"""CPU checks of parts of the DP refine that GPU runs only sample: the lean
kernels' schedules, modelled lane by lane against the oracle
(tests/emu/dp_lean_schedule.cpp, unbanded; tests/emu/dp_lean_ranged_schedule.cpp,
banded DTW and windowed LCSS: lane stagger, pair fetch and start, strip table,
halo and ring timing, result countdown, staged operands' provenance, LDS
bounds), the neighbour fixup that makes v_sqrt_f32 a correctly rounded
square root (tests/emu/sqrt_fixup.c), and the threshold compaction's
squared shortcut (tests/emu/compact_shortcut.c)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_lean_dp_schedule_model_matches_oracle(tmp_path):
    """k_dp_lean's schedule, with early abandoning of DTW and Frechet pairs
    against thresholds around their values (DTW also by the row minimum plus
    the suffix table of later rows' box bounds): every kept value equals the
    oracle's, no pair at or below its threshold is dropped, and pairs are
    dropped. Testing the row minimum at the row's padded end instead of its
    last real chunk (where a lane of the next strip may already have written
    the pair's result) is caught, and so are a suffix test without its
    rounding discount and a table entry that counts the tested strip."""
    exe = str(tmp_path / "dp_lean_schedule")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-o", exe,
                    os.path.join(HERE, "emu", "dp_lean_schedule.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp")],
                   check=True, timeout=300)
    r = subprocess.run([exe, "12"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    drops = [int(line.split(",")[1].split()[0]) for line in r.stdout.splitlines() if "dropped early" in line]
    assert sum(drops) > 0
    sfx = [int(line.split(":")[1]) for line in r.stdout.splitlines() if line.startswith("suffix-bound drops")]
    assert sfx and sfx[0] > 0
    r = subprocess.run([exe, "12", "late"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    # the suffix test without its rounding discount, and a table that also
    # counts the tested strip's rows, drop pairs at their own value
    for plant in ("nodisc", "sfx1"):
        r = subprocess.run([exe, "12", plant], capture_output=True, text=True, timeout=300)
        assert r.returncode != 0, plant


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_lean_ranged_dp_schedule_model_matches_oracle(tmp_path):
    """k_dp_lean_r's schedule for banded DTW (with early abandoning against
    thresholds around the pairs' values) and windowed LCSS: values equal the
    oracle's, no pair at or below its threshold is dropped, pairs are
    dropped. A halo ring half the launch's size, and a row minimum tested at
    the strip's padded end instead of the row's last in-band chunk, are
    caught."""
    exe = str(tmp_path / "dp_lean_ranged_schedule")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-o", exe,
                    os.path.join(HERE, "emu", "dp_lean_ranged_schedule.cpp"),
                    os.path.join(ROOT, "oracle", "oracle.cpp")], check=True, timeout=300)
    r = subprocess.run([exe, "40"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    drops = [int(line.split(",")[1].split()[0]) for line in r.stdout.splitlines() if "dropped early" in line]
    assert sum(drops) > 0
    r = subprocess.run([exe, "20", "2"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0, r.stdout[-2000:]
    r = subprocess.run([exe, "20", "1", "late"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0, r.stdout[-2000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_multi_wave_dp_schedule_model_matches_oracle(tmp_path):
    """k_dp_mw (a pair over 2 or 4 waves, barriers every 8 steps, two LDS
    halo buffers): every halo read finds the right strip and chunk written
    before the barrier that precedes it, with the waves of a superstep run
    in either order; values equal the oracle's bit for bit; DTW and Frechet
    pairs dropped early (against thresholds around their values) exceed
    their threshold, and are dropped before their result is written. A lag
    of 70 steps between strips (the kernel uses 72, the least safe is 71) is
    caught."""
    exe = str(tmp_path / "dp_mw_schedule")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-o", exe,
                    os.path.join(HERE, "emu", "dp_mw_schedule.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp")],
                   check=True, timeout=300)
    r = subprocess.run([exe, "12"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    assert int(r.stdout.split("failures, ")[1].split()[0]) > 0  # pairs dropped early
    r = subprocess.run([exe, "3", "70"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0, r.stdout[-2000:]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_sqrt_neighbour_fixup_is_correctly_rounded(tmp_path):
    """The DP's square root (v_sqrt_f32's result moved by the clamped signs
    of two fma residuals) equals sqrtf for every float x in [2^-96, 2^127)
    and every start within 1 ulp of the root, and for x = 0."""
    exe = str(tmp_path / "sqrt_fixup")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-o", exe,
                    os.path.join(HERE, "emu", "sqrt_fixup.c"), "-lm"], check=True, timeout=120)
    r = subprocess.run([exe, "all"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and " 0 mismatches" in r.stdout, r.stdout + r.stderr
    assert int(r.stdout.split()[0]) == 3 * ((0x7f0 - 0x0f8) << 20)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_compaction_shortcut_rejects_only_what_the_bound_rejects(tmp_path):
    """The compaction's squared gap-term test (DTW and Frechet) never rejects
    a pair whose f32 bound is at most the threshold, over 40M random pairs
    with the threshold at the bound or a few ulps above it and the hardware
    square root one ulp either way; with the margin cut from 2^-16 to 2^-21
    (a plant) it does."""
    exe = str(tmp_path / "compact_shortcut")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "emu", "compact_shortcut.c"),
                    "-lm"], check=True, timeout=120)
    r = subprocess.run([exe, "20000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "40000000 pairs, 0 violations", r.stdout + r.stderr
    r = subprocess.run([exe, "20000000", "21"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and not r.stdout.strip().endswith(" 0 violations"), r.stdout
