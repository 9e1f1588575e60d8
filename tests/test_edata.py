"""Edge values in a shard's edge-data blocks (gchi_edata_read / gchi_edata_write,
host file I/O only, no GPU): the layout of src/api/chifilenames.hpp:91-120 and
src/shards/memoryshard.hpp:268-300 -- <edata>.size holding the byte count and
blocks <edata>_blockdir_<B>/<id> of B = 1 MiB rounded up to a multiple of the
value size (the last one shorter), each one zlib stream when compressed
(src/util/ioutil.hpp:144-265).  Pinned against Python's zlib in both
directions, for raw blocks, for several blocks, and for the errors."""
import os
import zlib

import numpy as np
import pytest

import gchi

ERR_FORMAT = -5


def block_bytes(vb):
    b = 1 << 20
    while b % vb:
        b += 1
    return b


def write_reference_layout(ed, data, vb, compressed):
    """The blocks as the reference's sharder writes them (deflate level 1)."""
    b = block_bytes(vb)
    d = f"{ed}_blockdir_{b}"
    os.makedirs(d, exist_ok=True)
    with open(ed + ".size", "w") as f:
        f.write(str(len(data)))
    for i in range(0, max(len(data), 1), b):
        if i >= len(data):
            break
        chunk = data[i:i + b]
        with open(os.path.join(d, str(i // b)), "wb") as f:
            f.write(zlib.compress(chunk, 1) if compressed else chunk)


@pytest.mark.parametrize("compressed", [1, 0])
@pytest.mark.parametrize("vb,count", [(4, 1000), (4, 700_000), (8, 300_000), (12, 100_003)])
def test_edata_read_reference_layout(tmp_path, compressed, vb, count):
    rng = np.random.default_rng(vb * 7 + count)
    data = rng.integers(0, 256, vb * count, dtype=np.uint8).tobytes()
    ed = str(tmp_path / "g.edata..Z.e4B.0_1")
    write_reference_layout(ed, data, vb, compressed)
    out = np.zeros(vb * count, np.uint8)
    assert gchi.lib().gchi_edata_read(ed.encode(), compressed, vb, count, out.ctypes.data) == 0
    assert out.tobytes() == data


@pytest.mark.parametrize("compressed", [1, 0])
def test_edata_write_then_zlib_and_read(tmp_path, compressed):
    vb, count = 4, 600_000  # three blocks of 1 MiB, the last one shorter
    vals = np.arange(count, dtype=np.uint32) * np.uint32(2654435761)
    ed = str(tmp_path / "g.edata..Z.e4B.1_3")
    assert gchi.lib().gchi_edata_write(ed.encode(), compressed, vb, count, vals.ctypes.data) == 0
    assert open(ed + ".size").read().strip() == str(4 * count)
    d = f"{ed}_blockdir_{block_bytes(vb)}"
    assert sorted(os.listdir(d), key=int) == ["0", "1", "2"]
    raw = b""
    for i in range(3):
        blk = open(os.path.join(d, str(i)), "rb").read()
        raw += zlib.decompress(blk) if compressed else blk
    assert raw == vals.tobytes()
    back = np.zeros(count, np.uint32)
    assert gchi.lib().gchi_edata_read(ed.encode(), compressed, vb, count, back.ctypes.data) == 0
    np.testing.assert_array_equal(back, vals)


def test_edata_read_refuses_mismatches(tmp_path):
    vb, count = 4, 1000
    data = bytes(range(256)) * 15 + bytes(160)
    ed = str(tmp_path / "g.edata..Z.e4B.0_1")
    write_reference_layout(ed, data[:vb * count], vb, True)
    out = np.zeros(vb * (count + 1), np.uint8)
    L = gchi.lib()
    assert L.gchi_edata_read(ed.encode(), 1, vb, count + 1, out.ctypes.data) == ERR_FORMAT  # .size disagrees
    assert L.gchi_edata_read(ed.encode(), 0, vb, count, out.ctypes.data) == ERR_FORMAT  # compressed read as raw
    with open(ed + ".size", "w") as f:
        f.write(str(vb * count))
    with open(f"{ed}_blockdir_{block_bytes(vb)}/0", "wb") as f:
        f.write(b"not a zlib stream")
    assert L.gchi_edata_read(ed.encode(), 1, vb, count, out.ctypes.data) == ERR_FORMAT
    assert L.gchi_edata_read(str(tmp_path / "missing").encode(), 1, vb, count, out.ctypes.data) != 0
