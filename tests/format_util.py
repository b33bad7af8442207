"""Writers of the same reads in other input formats (test helpers): FASTQ
records and multi-line FASTA wrapped at a fixed width, with '\\n' or '\\r\\n'
line ends, from strict single-line FASTA."""
from __future__ import annotations


def records(fasta: bytes) -> list[tuple[bytes, bytes]]:
    """(header without '>', sequence) of every record of a strict single-line
    FASTA that starts with '>' and ends with '\\n'."""
    lines = fasta.split(b"\n")
    assert lines[-1] == b"" and len(lines) % 2 == 1
    return [(lines[i][1:], lines[i + 1]) for i in range(0, len(lines) - 1, 2)]


def to_fastq(fasta: bytes, eol: bytes = b"\n") -> bytes:
    out = []
    for h, s in records(fasta):
        out.append(b"@" + h + eol + s + eol + b"+" + eol + b"I" * len(s) + eol)
    return b"".join(out)


def to_multiline(fasta: bytes, width: int, eol: bytes = b"\n") -> bytes:
    out = []
    for h, s in records(fasta):
        out.append(b">" + h + eol)
        for i in range(0, len(s), width):
            out.append(s[i:i + width] + eol)
    return b"".join(out)
