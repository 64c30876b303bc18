"""The FASTQ reader (gis_fastq_*): chunks of records from one or two files,
plain or gzip, checked against Python's own line splitting with the
reference's record rules (four lines, read = line 2 up to '\\n', name = line 1
without '@' up to the first space). Host code only: no GPU is touched."""
import gzip
import os

import numpy as np
import pytest

import gpu_index_test_amd as G


def py_parse(data):
    """Independent parse: split at '\\n', drop the empty lines after the last
    record (keeping an empty last quality line), group by four."""
    lines = data.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    k = len(lines)
    while k and lines[k - 1] in (b"", b"\r"):
        k -= 1
    if k < len(lines) and k % 4 == 3:
        k += 1
    assert k % 4 == 0
    recs = []
    for i in range(0, k, 4):
        name = lines[i]
        sp = name.find(b" ")
        nm = name[1:sp] if sp >= 0 else name[1:]
        recs.append((nm, lines[i + 1]))
    return recs


def make_fastq(rng, n, lens=(0, 1, 50, 150, 151, 300), crlf=False, final_newline=True, trailing_blank=0):
    eol = b"\r\n" if crlf else b"\n"
    out = []
    for i in range(n):
        L = int(rng.choice(lens))
        seq = rng.choice(list(b"ACGTNacgtn"), size=L, p=[0.215] * 4 + [0.02] + [0.02] * 4 + [0.04])
        seq = seq.astype(np.uint8).tobytes()
        name = b"@read_%d" % i
        if i % 3 == 0:
            name += b" comment=%d extra" % (i * 7)
        if i % 5 == 0:
            name += b"\tTAB"
        qual = bytes(rng.integers(33, 74, size=L).astype(np.uint8))
        out.append(name + eol + seq + eol + b"+" + (b"read_%d" % i if i % 4 == 0 else b"") + eol + qual + eol)
    data = b"".join(out)
    if not final_newline and data.endswith(eol):
        data = data[: -len(eol)]
    return data + eol * trailing_blank


def collect(reader):
    reads, names, chunks = [], [], []
    for c in reader:
        chunks.append((c.index, c.first_record, c.n_records))
        reads.append([c.read(i) for i in range(c.n_reads)])
        names.append([c.name(i) for i in range(c.n_reads)])
    return reads, names, chunks


def write(tmp_path, name, data, gz=False):
    p = tmp_path / name
    if gz:
        with gzip.open(p, "wb", compresslevel=3) as f:
            f.write(data)
    else:
        p.write_bytes(data)
    return str(p)


@pytest.mark.parametrize("chunk_bytes,threads", [(4096, 1), (4096, 4), (65536, 3), (0, 0)])
@pytest.mark.parametrize("variant", ["plain", "crlf", "no_final_newline", "trailing_blank", "gzip"])
def test_single_end_matches_python(tmp_path, chunk_bytes, threads, variant):
    rng = np.random.default_rng(len(variant) * 31 + chunk_bytes % 97)
    data = make_fastq(rng, 700, crlf=variant == "crlf", final_newline=variant != "no_final_newline",
                      trailing_blank=3 if variant == "trailing_blank" else 0)
    path = write(tmp_path, "r.fq" + (".gz" if variant == "gzip" else ""), data, gz=variant == "gzip")
    with G.FastqReader(path, chunk_bytes=chunk_bytes, threads=threads) as r:
        reads, names, chunks = collect(r)
    exp = py_parse(data)
    assert [x for c in reads for x in c] == [s for _, s in exp]
    assert [x for c in names for x in c] == [n for n, _ in exp]
    # chunks in order, records counted from the start of the file
    first = 0
    for i, (idx, fr, n) in enumerate(chunks):
        assert idx == i and fr == first
        first += n
    if chunk_bytes == 4096:
        assert len(chunks) > 10


def test_paired_layout_and_equal_record_counts(tmp_path):
    rng = np.random.default_rng(3)
    d1 = make_fastq(rng, 900, lens=(150,))
    d2 = make_fastq(rng, 900, lens=(20, 150, 600))   # mate records of other sizes
    p1, p2 = write(tmp_path, "a_1.fq", d1), write(tmp_path, "a_2.fq.gz", d2, gz=True)
    with G.FastqReader(p1, p2, chunk_bytes=8192, threads=2) as r:
        m1, m2 = [], []
        for c in r:
            assert c.n_reads == 2 * c.n_records
            m1 += [c.read(i) for i in range(c.n_records)]
            m2 += [c.read(c.n_records + i) for i in range(c.n_records)]
    assert m1 == [s for _, s in py_parse(d1)]
    assert m2 == [s for _, s in py_parse(d2)]


def test_reads_longer_than_a_chunk(tmp_path):
    rng = np.random.default_rng(4)
    data = make_fastq(rng, 40, lens=(9000, 20000, 65535))
    path = write(tmp_path, "long.fq", data)
    with G.FastqReader(path, chunk_bytes=4096) as r:
        reads, _, _ = collect(r)
    assert [x for c in reads for x in c] == [s for _, s in py_parse(data)]


def test_empty_read_and_empty_file(tmp_path):
    data = b"@a\nACGT\n+\nIIII\n@b\n\n+\n\n"
    with G.FastqReader(write(tmp_path, "e.fq", data)) as r:
        reads, names, _ = collect(r)
    assert [x for c in reads for x in c] == [b"ACGT", b""]
    assert [x for c in names for x in c] == [b"a", b"b"]
    with G.FastqReader(write(tmp_path, "z.fq", b"")) as r:
        assert r.next_chunk() is None
    with G.FastqReader(write(tmp_path, "z2.fq", b"\n\n")) as r:
        assert r.next_chunk() is None


@pytest.mark.parametrize("data,what", [
    (b"@a\nACGT\n+\nIIII\nb\nACGT\n+\nIIII\n", "'@'"),
    (b"@a\nACGT\n-\nIIII\n", "'+'"),
    (b"@a\nACGT\n+\nIIII\n@b\nACGT\n", "incomplete"),
    (b"@a\n" + b"A" * 65536 + b"\n+\n" + b"I" * 65536 + b"\n", "longer"),
])
def test_malformed_input_is_reported(tmp_path, data, what):
    with G.FastqReader(write(tmp_path, "bad.fq", data), chunk_bytes=4096) as r:
        with pytest.raises(G.GisError) as e:
            for _ in r:
                pass
    assert e.value.code == 11 and what in str(e.value)


def test_mates_with_different_counts(tmp_path):
    rng = np.random.default_rng(5)
    p1 = write(tmp_path, "m1.fq", make_fastq(rng, 30, lens=(100,)))
    p2 = write(tmp_path, "m2.fq", make_fastq(rng, 31, lens=(100,)))
    with G.FastqReader(p1, p2) as r:
        with pytest.raises(G.GisError) as e:
            for _ in r:
                pass
    assert e.value.code == 11 and "different numbers of records" in str(e.value)


def test_missing_file(tmp_path):
    with pytest.raises(G.GisError) as e:
        G.FastqReader(str(tmp_path / "nope.fq"))
    assert e.value.code == 7


def test_good_chunks_come_before_the_error(tmp_path):
    rng = np.random.default_rng(6)
    good = make_fastq(rng, 200, lens=(150,))
    path = write(tmp_path, "late.fq", good + b"@x\nACGT\n")
    got = 0
    with G.FastqReader(path, chunk_bytes=4096) as r:
        with pytest.raises(G.GisError):
            for c in r:
                got += c.n_records
    assert got == 200
    assert os.path.getsize(path) > 4096
