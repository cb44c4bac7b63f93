"""Write the pair-selection golden vectors with the reference's own Perl scripts.

Runs ONLY in a container that holds the reference tree (the scripts are read
from /root/reference and executed with the system perl); the GPU box and the
tests only read the committed outputs under tests/golden/nbkc/.

Inputs: seeded per-read integer stats {median, nwin, sum, sumsq} (the
aggregates a coverage pass produces), written as two C-locale-sorted stats
files in the column layout `fastaToKmerCoverageStats` prints here
(acc median_cov mean_cov stdev pct_stdev_of_avg).  The pairs cover:
  * accessions with /1 and /2 suffixes of varied length (so `lt` order and
    the C-locale sort interleave), and left-only / right-only accessions
    (skipped by `nbkc_merge_left_right_stats.pl --sorted`, :113-125);
  * reads with no window (median 0, mean 0: merged mean "0.0" is discarded,
    nbkc_normalize.pl:99-102) on one or both sides;
  * medians below min_cov (nbkc_normalize.pl:91-94);
  * coefficients of variation around 1 and 2 (the max_CV test, :104-109).
Outputs: the merged pairs file and, for each (max_cov, min_cov, max_CV), the
selected accessions (stdout) and the three summary lines (stderr) of
`nbkc_normalize.pl`; plus the single-end run of nbkc_normalize.pl on the left
file (insilico_read_normalization.pl:917-944).

    python tests/golden/make_nbkc_fixtures.py
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402  (formats rows exactly as the CLI does)

REF = Path("/root/reference/util/support_scripts")
OUT = HERE / "nbkc"
SETTINGS = [(50, 0, 10000), (50, 1, 10000), (200, 0, 10000), (5, 3, 10000), (50, 0, 1), (30, 10, 2)]
SEED = 20261203
N_PAIRS = 1200


def read_stats(rng: np.random.Generator, n: int) -> np.ndarray:
    st = np.zeros(n, O.READ_STAT_DTYPE)
    for i in range(n):
        u = rng.random()
        nwin = 0 if u < 0.06 else int(rng.integers(1, 127))
        if nwin == 0:
            continue
        level = float(np.exp(rng.normal(2.5, 1.6)))
        if rng.random() < 0.15:  # a read over an error-rich or low region
            level = float(rng.uniform(0.5, 4))
        v = np.maximum(1, rng.poisson(level, nwin)).astype(np.int64)
        if rng.random() < 0.2:  # a run of missing k-mers (counted as 1)
            a = int(rng.integers(0, nwin))
            v[a:a + int(rng.integers(1, 30))] = 1
        s = np.sort(v)
        st[i]["median"] = int(s[nwin // 2])
        st[i]["nwin"] = nwin
        st[i]["sum"] = int(v.sum())
        sq = int((v * v).sum())
        st[i]["sumsq_lo"] = sq & (2**64 - 1)
        st[i]["sumsq_hi"] = sq >> 64
    return st


def main() -> None:
    rng = np.random.default_rng(SEED)
    OUT.mkdir(exist_ok=True)
    ids = rng.choice(np.arange(1, 3_000_000), size=N_PAIRS + 80, replace=False)
    names = [f"s{x}" for x in ids]
    left_only = set(names[N_PAIRS:N_PAIRS + 40])
    right_only = set(names[N_PAIRS + 40:])
    left_names = sorted([n for n in names if n not in right_only], key=lambda s: (s + "/1").encode())
    right_names = sorted([n for n in names if n not in left_only], key=lambda s: (s + "/2").encode())
    stl = read_stats(rng, len(left_names))
    str_ = read_stats(rng, len(right_names))
    # a few pairs with no window on both sides, and a few on one side
    both = rng.choice(len(left_names), 12, replace=False)
    for i in both:
        nm = left_names[i]
        if nm in right_names:
            stl[i] = 0
            str_[right_names.index(nm)] = 0
    rows = []
    for fn, nm_list, st, suf in (("left.stats.sort", left_names, stl, "/1"),
                                 ("right.stats.sort", right_names, str_, "/2")):
        with open(OUT / fn, "w") as f:
            f.write("acc\tmedian_cov\tmean_cov\tstdev\tpct_stdev_of_avg\n")
            for nm, s in zip(nm_list, st):
                f.write(O.format_stats_row(nm + suf, s))
        # the integer aggregates behind each row (GPU selection input)
        with open(OUT / fn.replace(".stats.sort", ".aggregates.tsv"), "w") as f:
            f.write("acc\tmedian\tnwin\tsum\tsumsq_lo\tsumsq_hi\n")
            for nm, s in zip(nm_list, st):
                f.write(f"{nm}{suf}\t{s['median']}\t{s['nwin']}\t{s['sum']}\t{s['sumsq_lo']}\t{s['sumsq_hi']}\n")
        rows.append(len(nm_list))

    def perl(script, args, stdout_path, stderr_path=None):
        r = subprocess.run(["perl", str(REF / script), *args], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, check=True, cwd=OUT)
        (OUT / stdout_path).write_bytes(r.stdout)
        if stderr_path:
            (OUT / stderr_path).write_bytes(r.stderr)

    perl("nbkc_merge_left_right_stats.pl",
         ["--left", "left.stats.sort", "--right", "right.stats.sort", "--sorted"], "pairs.stats")
    for mx, mn, cv in SETTINGS:
        tag = f"C{mx}.minC{mn}.maxCV{cv}"
        perl("nbkc_normalize.pl", ["--stats_file", "pairs.stats", "--max_cov", str(mx), "--min_cov", str(mn),
                                   "--max_CV", str(cv)], f"pairs.{tag}.accs", f"pairs.{tag}.stderr")
    perl("nbkc_normalize.pl", ["--stats_file", "left.stats.sort", "--max_cov", "50", "--min_cov", "0",
                               "--max_CV", "10000"], "left.C50.minC0.maxCV10000.accs",
         "left.C50.minC0.maxCV10000.stderr")
    ver = subprocess.run(["perl", "-e", "print $^V"], stdout=subprocess.PIPE, text=True, check=True).stdout
    (OUT / "README.txt").write_text(
        "Golden vectors of the pair merge and normalization selection, written by\n"
        "tests/golden/make_nbkc_fixtures.py with the reference's\n"
        "util/support_scripts/nbkc_merge_left_right_stats.pl --sorted and nbkc_normalize.pl\n"
        f"under perl {ver} (seed {SEED}; {rows[0]} left rows, {rows[1]} right rows).\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
