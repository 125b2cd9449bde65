"""bench.py without a GPU: the one-GPU guard and the algorithmic byte counts
the roofline divides by (each restated here from its parts)."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_headline_workload_shape():
    assert (bench.M, bench.G, bench.T, bench.L, bench.S) == (512, 4096, 4, 32, 3)
    assert bench.M * bench.G == 2097152


def test_cross_step_bytes():
    M, G, T, S = bench.M, bench.G, bench.T, bench.S
    # per particle: resampling reads the sorted log-weight and state, the
    # sampling reads 3 uniforms, the sort writes the sorted log-weight and state
    assert bench.XS_BYTES_PER_PARTICLE == 4 + 4 * S + 4 * 3 + 4 + 4 * S
    # per sub-filter: offset and two gates; four neighbours' t received
    # particles read and their weights written; the argmax's winner and
    # partials; the sort's winner, partials and best t
    recv = 4 * T * (1 + S) * 4 + 4 * T * 4
    argmax = 8 + 4 * (2 + S)
    sort = 8 + 4 * (2 + S) + 4 * T * (1 + S)
    assert bench.XS_BYTES_PER_SUB == 12 + recv + argmax + sort == 452
    assert bench.XS_BYTES_PER_LAUNCH == 44 * M * G + 452 * G == 94126080


def test_sample_sort_and_mtgp_bytes():
    M, G, T, S = bench.M, bench.G, bench.T, bench.S
    assert bench.SS_BYTES_PER_SUB == 4 + 8 + 4 * (2 + S) + 4 * T * (1 + S) == 96
    assert bench.SW_BYTES_PER_LAUNCH == 40 * M * G
    # the buffer of ceil((3n + n + g) / 73728) * 73728 words, plus the 96
    # streams' statuses read and written (351 words each)
    n = M * G
    words = -(-(3 * n + n + G) // 73728) * 73728
    assert bench.RNG_BYTES_PER_LAUNCH == 4 * words + 2 * 4 * 351 * 96


def test_dh6_bytes():
    assert bench.DH_BYTES_PER_PARTICLE == 48 + 24 + 48 + 4
    assert bench.DH_BYTES_PER_SUB == 4 + 8 + 4 * 14 + 4 * 13 == 120
