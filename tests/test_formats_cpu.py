"""The CPU restatement of the input-format rewrite (oracle.normalize) against
known answers and against invariants that fix it: strict single-line FASTA
is left as it is, and the same reads written as FASTQ or as wrapped
multi-line FASTA, with '\\n' or '\\r\\n' line ends, rewrite to the original."""
from __future__ import annotations

import pytest

from format_util import to_fastq, to_multiline
from golden_util import read_fasta

KNOWN = [
    ("fasta", b"", b""),
    ("fasta", b">a\nAC\nGT\n>b\n>c\nTT", b">a\nACGT\n>b\n\n>c\nTT\n"),
    ("fasta", b"junk\n\n>a\r\nAC\r\n\r\nGT\r\n", b">a\nACGT\n"),
    ("fasta", b"no record at all\nACGT\n", b""),
    ("fasta", b">only header", b">only header\n\n"),
    ("fasta", b">a\nAC>GT\n", b">a\nAC>GT\n"),
    ("fasta", b">a\nAC\rGT\r\r\n", b">a\nAC\rGT\r\n"),
    ("fastq", b"@r1\nACGT\n+\nIIII\n@r2\r\nGG\r\n+r2\r\n!!\r\n", b">r1\nACGT\n>r2\nGG\n"),
    ("fastq", b"@r1\nACGT\n+\n@III", b">r1\nACGT\n"),
    ("fastq", b"@\n\n+\n\n", b">\n\n"),
]


@pytest.mark.parametrize("fmt,data,want", KNOWN)
def test_normalize_known_answers(oracle_mod, fmt, data, want):
    assert oracle_mod.normalize(data, fmt) == want


@pytest.mark.parametrize("data", [b"@r\nAC\n+\n", b">r\nAC\n+\nII\n", b"@r\nAC\n-\nII\n", b"@r\nAC\n+\nII\n\n"])
def test_normalize_fastq_errors(oracle_mod, data):
    with pytest.raises(ValueError):
        oracle_mod.normalize(data, "fastq")


@pytest.mark.parametrize("name", ["syn_tiny.fa.gz", "syn_mid.fa.gz", "read_test1.fa", "ref_test_two_reads.fa"])
def test_normalize_round_trips(oracle_mod, name):
    fa = read_fasta(name)
    assert fa.startswith(b">") and fa.endswith(b"\n")
    assert oracle_mod.normalize(fa, "fasta") == fa
    for eol in (b"\n", b"\r\n"):
        assert oracle_mod.normalize(to_fastq(fa, eol), "fastq") == fa
        for width in (1, 7, 60):
            assert oracle_mod.normalize(to_multiline(fa, width, eol), "fasta") == fa

