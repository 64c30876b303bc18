"""FASTQ ingestion into the serving queue: every chunk of a FASTQ file (one
file or two mates files, plain or gzip) submitted as a request returns the
oracle's NAMs for the reads the file holds, and the pair step on a paired
chunk returns the oracle's pair results."""
import gzip

import numpy as np
import pytest

import gpu_index_test_amd as G
from oracle import oracle_lib as O
from tests.test_fastq import py_parse

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def genome_index(ctx):
    g, lens = G.synth_genome(300_000, n_contigs=2, gc=0.43, repeat_fraction=0.08, n_families=10,
                             family_len=(300, 2000), copies=(2, 25), divergence=(0.0, 0.02),
                             tandem_fraction=0.02, tandem_unit=(1, 10), seed=71)
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(None, p)
    idx.populate_arrays(g, lens, f=0.0002, n_threads=8)
    oidx = O.Index.build_arrays(g, lens, O.params_from_read_length(150), n_threads=8)
    return g, lens, idx, oidx


def _fastq_bytes(reads, tag, crlf=False):
    eol = b"\r\n" if crlf else b"\n"
    out = []
    for i, r in enumerate(reads):
        out.append(b"@%s_%d extra words%s%s%s+%s%s%s" % (tag, i, eol, r, eol, eol, b"I" * len(r), eol))
    return b"".join(out)


def _write(path, data, gz):
    if gz:
        with gzip.open(path, "wb", compresslevel=2) as f:
            f.write(data)
    else:
        path.write_bytes(data)
    return str(path)


@pytest.mark.parametrize("paired,gz", [(False, False), (True, False), (True, True)])
def test_fastq_requests_match_oracle(genome_index, tmp_path, paired, gz):
    g, lens, idx, oidx = genome_index
    n = 2500
    s, o, l, _ = G.synth_reads(g, lens, n, lengths=(100, 150, 250), weights=(0.3, 0.5, 0.2),
                               paired=paired, sub_rate=0.01, indel_rate=0.001, seed=72 + paired)
    reads = [s[int(o[i]): int(o[i]) + int(l[i])].tobytes() for i in range(len(l))]
    if paired:
        m1, m2 = reads[0::2], reads[1::2]
        d1, d2 = _fastq_bytes(m1, b"p"), _fastq_bytes(m2, b"p", crlf=False)
        p1 = _write(tmp_path / ("r1.fq" + (".gz" if gz else "")), d1, gz)
        p2 = _write(tmp_path / ("r2.fq" + (".gz" if gz else "")), d2, gz)
        reader = G.FastqReader(p1, p2, chunk_bytes=64 << 10)
    else:
        d1 = _fastq_bytes(reads, b"s")
        reader = G.FastqReader(_write(tmp_path / "r.fq", d1, False), chunk_bytes=64 << 10)
    srv = G.Server(idx, slots=3, max_step_reads=4096)
    try:
        # keep up to 6 chunks in flight: the reader and the server overlap
        pending, results = [], []
        while True:
            rq, chunk = srv.submit_fastq(reader)
            if rq is not None:
                pending.append((rq, chunk))
            if len(pending) >= 6 or (rq is None and pending):
                rq0, ch0 = pending.pop(0)
                res, _ = srv.wait(rq0)
                results.append((ch0, res))
            if rq is None and not pending:
                break
        assert len(results) > 3
        seen = []
        for ch, res in results:
            cr = [ch.read(i) for i in range(ch.n_reads)]
            seen += cr[: ch.n_records]
            seqs, offs, ls = G.pack_reads(cr)
            eo, en, ef, efl = O.map_batch(oidx, seqs, offs, ls, 2 * oidx.filter_cutoff, 1000, 8)
            assert np.array_equal(res.nam_offsets, eo)
            assert res.nams.tobytes() == en.tobytes()
            if paired:
                # the chunk's layout is the pair step's: mate 1 block, mate 2 block
                m = G.Mapper(idx, len(ls), int(ls.sum()))
                m.upload(seqs, offs, ls)
                m.run()
                pr = m.pairs()
                exp = O.pairs_batch(eo, en, ch.n_records)
                assert pr.nams.tobytes() == exp["ord"].tobytes()
                assert np.array_equal(pr.types, exp["types"])
                assert pr.joint.tobytes() == exp["joint"].tobytes()
                m.close()
        assert seen == [x for _, x in py_parse(d1)]
    finally:
        srv.close()
        reader.close()
