"""Write FASTQ->FASTA golden vectors with the reference's fastQ_to_fastA.pl.

Runs ONLY in a container that holds the reference tree (the script is read
from /root/reference and executed with the system perl); the tests only read
the committed outputs under tests/golden/ingest/.

Inputs: two small FASTQ files written here with the cases the script's rules
branch on (util/support_scripts/fastQ_to_fastA.pl:66-117):
  * names with and without /1,/2 suffixes, Casava 1.8 "1:N:0:..." and
    "2:Y:..." fields, a field that is not Casava ("3:N", "x1:"), tabs;
  * \\r\\n line endings (only the header line loses its \\r), a header that
    starts with a blank after '@', '+name' quality headers;
  * lowercase and IUPAC bases (for --rev), blank lines between records;
  * edge_dirty.fq adds malformed records (length mismatch, missing '+',
    blank quality), fatal without --ignore_dirty.
Outputs: stdout of each option set, plus the SHA-256 of the script's output
on the sample reads (tests/golden/sample_data) with --rev and with -a 2.

    python tests/golden/make_ingest_fixtures.py
"""
from __future__ import annotations

import hashlib
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "ingest"
SCRIPT = Path("/root/reference/util/support_scripts/fastQ_to_fastA.pl")

RECORDS = [
    ("@r1/1", "ACGTACGTNN", "+", "IIIIIIIIII"),
    ("@r2 1:N:0:ATCACG", "acgtRYKMUu", "+r2", "IIIIIIIIII"),
    ("@r3 2:Y:18:GGCTAC extra", "GGGGCCCCAT", "+", "##########"),
    ("@r4/2 1:N:0:X", "TTTTAAAACG", "+", "IIIIIIIIII"),
    ("@r5 3:N:0:X", "CAGTCAGTCA", "+", "IIIIIIIIII"),
    ("@r6 x1:N", "GATTACAGAT", "+", "IIIIIIIIII"),
    ("@r7\t1:N:0:TAB", "AAAAACCCCC", "+", "IIIIIIIIII"),
    ("@ r8 1:N", "CCCCCAAAAA", "+", "IIIIIIIIII"),
    ("@r9_forward 1:N:0:A", "ACGTTGCAAC", "+", "IIIIIIIIII"),
    ("@r10/1", "N", "+", "I"),
]


def write_fq(path: Path, dirty: bool) -> None:
    lines = []
    for i, (h, s, p, q) in enumerate(RECORDS):
        eol = "\r\n" if i % 3 == 1 else "\n"
        lines.append(h + eol + s + "\n" + p + "\n" + q + "\n")
        if i == 4:
            lines.append("\n")  # a blank line between records
    if dirty:
        lines.insert(3, "@bad1 1:N\nACGTA\n+\nIIII\n")        # quality shorter than sequence
        lines.insert(6, "@bad2\nACGTA\nIIIII\nIIIII\n")       # no '+' line
        lines.insert(9, "@bad3\nACGTA\n+\n     \n")           # blank quality
    path.write_bytes("".join(lines).encode())


def main() -> None:
    OUT.mkdir(exist_ok=True)
    write_fq(OUT / "edge.fq", False)
    write_fq(OUT / "edge_dirty.fq", True)

    def perl(args, out_name):
        r = subprocess.run(["perl", str(SCRIPT), *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           cwd=OUT)
        (OUT / out_name).write_bytes(r.stdout)
        return r.returncode

    assert perl(["-I", "edge.fq"], "edge.fa") == 0
    assert perl(["-I", "edge.fq", "-a", "1"], "edge.a1.fa") == 0
    assert perl(["-I", "edge.fq", "--rev"], "edge.rev.fa") == 0
    assert perl(["-I", "edge.fq, edge.fq"], "edge.twice.fa") == 0
    assert perl(["-I", "edge_dirty.fq", "--ignore_dirty"], "edge_dirty.ignore.fa") == 0
    rc = perl(["-I", "edge_dirty.fq"], "edge_dirty.fatal.fa")
    assert rc != 0
    (OUT / "edge_dirty.fatal.rc").write_text(f"{rc}\n")
    sample = HERE / "sample_data"
    sums = []
    for f, args in (("reads.left.fq.gz", ["--rev"]), ("reads.right.fq.gz", ["-a", "2"])):
        r = subprocess.run(["perl", str(SCRIPT), "-I", str(sample / f), *args], stdout=subprocess.PIPE, check=True)
        sums.append(f"{hashlib.sha256(r.stdout).hexdigest()}  {f} {' '.join(args)}\n")
    (OUT / "sample_sha256.txt").write_text("".join(sums))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
