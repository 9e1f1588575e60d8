"""GPU parity of the device decode of shard adjacency streams (.adj bytes):
edge pairs bit-exact vs the oracle's restatement of memory_shard's parse, on
the reference sharder's own bytes (golden cases), on full-size shards written
by the device sharder, and on streams whose records and zero runs cross many
32 KB chunks; truncated streams are refused."""
import numpy as np
import pytest

import graphs
import oracle
import test_oracle

pytestmark = pytest.mark.gpu


def to_dev_u8(dev, b):
    torch = dev._torch()
    a = np.frombuffer(bytes(b), np.uint8) if len(b) else np.zeros(1, np.uint8)
    return torch.from_numpy(a.copy()).to("cuda")


def decode(dev, b):
    s, d, vend = dev.shard_decode(to_dev_u8(dev, b), len(b))
    got = oracle.shard_decode(b)
    ne = len(got[0])
    return dev.to_host_u32(s, ne), dev.to_host_u32(d, ne), vend, got


@pytest.mark.parametrize("case", test_oracle.SHARDED, ids=[c["name"] for c in test_oracle.SHARDED])
def test_decode_reference_shard_bytes(dev, case, manifest):
    """The reference's .adj bytes (literal where the manifest holds them, else
    the oracle encoding whose SHA-256 equals the reference file's)."""
    import hashlib
    src, dst, n = graphs.make(case)
    m = manifest[case["name"]]["shards"]
    _, shards = test_oracle.reference_shards(src, dst, n, case.get("nshards", 1))
    for (s, d), g in zip(shards, m["shards"]):
        adj = bytes.fromhex(g["adj_hex"]) if g.get("adj_hex") else oracle.shard_encode(s, d)[0]
        assert hashlib.sha256(adj).hexdigest() == g["adj_sha"]
        gs, gd, vend, _ = decode(dev, adj)
        np.testing.assert_array_equal(gs, s)
        np.testing.assert_array_equal(gd, d)
        assert vend == (int(s[-1]) + 1 if len(s) else vend)


def test_decode_device_sharder_scale20(dev):
    """Every shard the device sharder writes for a scale-20 R-MAT graph decodes
    back to that shard's (src, dst) records."""
    src, dst = graphs.rmat_pairs(20, 16, 2)
    src, dst = graphs.preprocess(src, dst)
    maxid = int(max(src.max(), dst.max()))
    sh = dev.Shards(dev.to_dev_u32(src), dev.to_dev_u32(dst), maxid, 5)
    total = 0
    for p in range(sh.nshards):
        adj, _ = sh.encode(p)
        hs, hd = sh.edges_host(p)
        gs, gd, _, want = decode(dev, adj)
        np.testing.assert_array_equal(gs, want[0])
        np.testing.assert_array_equal(gd, want[1])
        np.testing.assert_array_equal(gs, hs)
        np.testing.assert_array_equal(gd, hd)
        total += len(gs)
    sh.close()
    assert total == len(src)


def test_decode_records_across_chunks(dev):
    """A first source past 5M (39 KB of zero runs before the first record), a
    70000-edge source (a 280 KB record over nine chunks), a gap of 300000
    vertices, and many 1-3 edge rows around chunk boundaries."""
    rng = np.random.default_rng(11)
    src = np.concatenate([np.full(2, 5_000_000), np.full(70000, 5_000_001), np.full(3, 5_300_002),
                          rng.integers(5_400_000, 5_500_000, 60000)]).astype(np.uint32)
    dst = rng.integers(0, 1 << 32, len(src), dtype=np.uint64).astype(np.uint32)
    o = np.lexsort((dst, src))
    src, dst = src[o], dst[o]
    adj, _ = oracle.shard_encode(src, dst)
    assert len(adj) > 8 * 32768
    gs, gd, vend, _ = decode(dev, adj)
    np.testing.assert_array_equal(gs, src)
    np.testing.assert_array_equal(gd, dst)
    assert vend == int(src[-1]) + 1
    # the stream cut inside a record (first zero run, a zero-run pair at a chunk
    # edge, the last dst word) is refused
    for cut in (1, 3, 32769, len(adj) - 1):
        with pytest.raises(dev.GchiError):
            dev.shard_decode(to_dev_u8(dev, adj[:cut]), cut)


def test_decode_small_and_malformed(dev):
    s, d, vend = dev.shard_decode(to_dev_u8(dev, b""), 0)
    assert vend == 0
    gs, gd, vend, _ = decode(dev, bytes.fromhex("00040200000000070000000101000000"))
    assert gs.tolist() == [5, 5, 6] and gd.tolist() == [0, 7, 1] and vend == 7
    for bad in (b"\x00", b"\xff\x01\x00", b"\x02\x01\x00\x00\x00", b"\x01\x00\x00\x00"):
        with pytest.raises(dev.GchiError):
            dev.shard_decode(to_dev_u8(dev, bad), len(bad))


def test_decode_many_streams_in_one_call(dev):
    """All shards of a scale-18 graph, plus an empty stream and a golden hand
    stream, decoded in one call: the edges land back to back in stream order,
    each stream's src ids counting from 0."""
    src, dst = graphs.rmat_pairs(18, 16, 5)
    src, dst = graphs.preprocess(src, dst)
    maxid = int(max(src.max(), dst.max()))
    sh = dev.Shards(dev.to_dev_u32(src), dev.to_dev_u32(dst), maxid, 4)
    streams = [sh.encode(p)[0] for p in range(sh.nshards)]
    sh.close()
    streams.insert(2, b"")
    streams.append(bytes.fromhex("00040200000000070000000101000000"))
    d = [(to_dev_u8(dev, b), len(b)) for b in streams]
    gs, gd, counts, vends = dev.shard_decode_many(d)
    want = [oracle.shard_decode(b) for b in streams]
    assert counts.tolist() == [len(w[0]) for w in want]
    tot = int(counts.sum())
    np.testing.assert_array_equal(dev.to_host_u32(gs, tot), np.concatenate([w[0] for w in want]))
    np.testing.assert_array_equal(dev.to_host_u32(gd, tot), np.concatenate([w[1] for w in want]))
    assert vends.tolist() == [int(w[0][-1]) + 1 if len(w[0]) else 0 for w in want]
    # one truncated stream among good ones fails the call
    bad = d[:2] + [(to_dev_u8(dev, streams[0][:-1]), len(streams[0]) - 1)]
    with pytest.raises(dev.GchiError):
        dev.shard_decode_many(bad)
