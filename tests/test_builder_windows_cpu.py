"""The restatement of the implicit builder counters' window count
(tests/builder_util.windows_closed_form; concurrent_count_and_process_omp_implicit_v0
/ _v1, src/concurrent_kmer_counter.hpp:910-1022,1363-1578) pinned by known
answers worked out by hand, at one thread: records, windows, distinct windows
under strncmp(a, b, klen) (raw bytes up to the first NUL), and the two error
cases.  The packed-key figures of the explicit counter (builder_closed_form)
get a hand-worked case here too; test_gpu_builder.py pins that restatement
against the oracle on larger inputs."""
from __future__ import annotations

import pytest

from builder_util import builder_closed_form, windows_closed_form


def test_two_equal_reads():
    # windows ACGT CGTA GTAC, twice
    assert windows_closed_form(b">a\nACGTAC\n>b\nACGTAC\n", 4, 6, "v0") == (2, 6, 3)


def test_case_and_other_bytes_are_distinct_windows():
    # as strings: ACG CGT GTN / acg cgt gtA, six different windows
    fa = b">a\nACGTN\n>b\nacgtA\n"
    assert windows_closed_form(fa, 3, 5, "v1") == (2, 6, 6)
    # packed (symbol_to_kmer: A C T G -> 0 1 2 3, anything else 0; the word
    # keeps every symbol since the read start): read a gives 0b000111 = 7,
    # 0b01_11_10 = 30, 0b01_11_10_00 = 120; read b is all zeros three times
    reads, counts = builder_closed_form(fa, 3, 5, 64)
    assert reads == 2 and dict(counts) == {7: 1, 30: 1, 120: 1, 0: 3}


def test_nul_ends_the_comparison():
    # strncmp stops at the first NUL: "AC\0G" == "AC\0T", "C\0GT" == "C\0TT"
    fa = b">a\nAC\x00GT\n>b\nAC\x00TT\n"
    assert windows_closed_form(fa, 4, 5, "v0") == (2, 4, 2)


def test_header_bytes_inside_windows():
    # read_len 8 > line length 4: windows run into the next header line
    fa = b">a\nACGT\n>b\nACGT\n"
    # read a = "ACGT\n>b\n" (8 bytes): 5 windows of 4; read b = "ACGT\n" (cut
    # at the end: 5 bytes, 2 windows)
    assert windows_closed_form(fa, 4, 8, "v1") == (2, 7, 5)  # ACGT CGT\n GT\n> T\n>b \n>b\n, ACGT CGT\n
    with pytest.raises(ValueError, match="Variable size read"):
        windows_closed_form(fa, 4, 8, "v0")


def test_short_reads():
    fa = b">a\nACGTAC\n>b\nACG"
    # v0: the last read is cut short by the end of the buffer
    with pytest.raises(ValueError):
        windows_closed_form(fa, 2, 6, "v0")
    # v1 takes it: windows AC CG GT TA AC, AC CG -> distinct AC CG GT TA
    assert windows_closed_form(fa, 2, 6, "v1") == (2, 7, 4)
    # ... unless it is shorter than klen
    with pytest.raises(ValueError):
        windows_closed_form(fa, 4, 6, "v1")


def test_records_follow_the_scan():
    # junk before the first '>', a second '>' inside a header, a header with
    # no delimiter at the end: two records
    fa = b"junk\n>h1>x\nAAAA\n>h2\nCCCC\n>tail"
    assert windows_closed_form(fa, 2, 4, "v0") == (2, 6, 2)
    assert windows_closed_form(b"", 2, 4, "v0") == (0, 0, 0)
