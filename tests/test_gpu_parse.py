"""GPU parity of the device edge-list parser (gchi_parse_edgelist) against the
oracle's restatement of the reference's convert_edgelist (whose atoi is the C
library's own): bit-exact pairs in file order, self-loops dropped, max id;
malformed lines refused."""
import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu


def check_same(dev, text):
    s, d, n, mx = dev.parse_edgelist(text)
    ws, wd = oracle.parse_edgelist(text)
    assert n == len(ws)
    np.testing.assert_array_equal(dev.to_host_u32(s, n), ws)
    np.testing.assert_array_equal(dev.to_host_u32(d, n), wd)
    assert mx == (int(max(ws.max(), wd.max())) if n else 0)


def test_parse_tricky_lines(dev):
    text = (b"# comment\n% another\n0\t1\n2,3\n  4   5  \n6\t7\t0.5\n8 8\n9\t10\r\n"
            b"+11,-1\n12abc 13x\n14 15")
    check_same(dev, text)
    check_same(dev, b"1 2\n")
    check_same(dev, b"# only a comment\n")
    check_same(dev, b"")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_parse_random_layouts(dev, seed):
    """Lines of random shape crossing the 64-byte lane slices and 4 KiB chunk
    boundaries at every offset: runs of delimiters, comment lines, values,
    CRLF, long (but under 1023-byte) lines."""
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(20000):
        k = rng.integers(0, 20)
        if k == 0:
            parts.append("#" + "x" * int(rng.integers(0, 300)) + "\n")
            continue
        if k == 1:
            parts.append("%c\n")
            continue
        a, b = int(rng.integers(0, 1 << 20)), int(rng.integers(0, 1 << 20))
        if k == 2:
            b = a
        d1 = "".join(rng.choice(["\t", " ", ","], int(rng.integers(1, 4))))
        pre = " " * int(rng.integers(0, 3))
        tail = "" if k > 5 else d1 + f"{rng.random():.6f}"
        pad = " " * int(rng.integers(0, 900)) if k == 3 else ""
        eol = "\r\n" if k == 4 else "\n"
        parts.append(f"{pre}{a}{d1}{b}{tail}{pad}{eol}")
    text = "".join(parts).encode()
    check_same(dev, text)
    check_same(dev, text.rstrip(b"\n"))  # no final newline


def test_parse_golden_inputs_and_rmat20(dev):
    """The edge-list files the golden fixtures were made from, and an R-MAT
    scale-20 list (16M lines, 4 KiB chunks by the thousand)."""
    for case in graphs.PR_CASES[:3]:
        src, dst, n = graphs.make(case)
        txt = "".join(f"{a}\t{b}\n" for a, b in zip(src.tolist(), dst.tolist())).encode()
        s, d, ne, mx = dev.parse_edgelist(txt)
        np.testing.assert_array_equal(dev.to_host_u32(s, ne), src)
        np.testing.assert_array_equal(dev.to_host_u32(d, ne), dst)
    hs, hd = dev.rmat_host(20, 16 << 20, 3)
    txt = ("\n".join(f"{a} {b}" for a, b in zip(hs.tolist(), hd.tolist())) + "\n").encode()
    s, d, ne, mx = dev.parse_edgelist(txt)
    keep = hs != hd
    assert ne == int(keep.sum())
    np.testing.assert_array_equal(dev.to_host_u32(s, ne), hs[keep])
    np.testing.assert_array_equal(dev.to_host_u32(d, ne), hd[keep])
    assert mx == int(max(hs[keep].max(), hd[keep].max()))


@pytest.mark.parametrize("bad", [b"1 2\n\n3 4\n", b"7\n", b" , \n", b"1 2\n" + b"3 " * 600 + b"\n"],
                         ids=["blank", "one_field", "delims_only", "too_long"])
def test_parse_refuses_malformed(dev, bad):
    with pytest.raises(dev.GchiError, match="FORMAT|format|fields"):
        dev.parse_edgelist(bad)
