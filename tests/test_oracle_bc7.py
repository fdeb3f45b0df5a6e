"""Pins for the oracle's BC7 decoder (oracle/bc7.cpp), the format of the reference's images
(Packager.cpp:216-308, Graphics.cpp:3620-3650).

The reference ships no BC7 files and NVTT is closed, so the pins come from the format itself and
from an independent decoder:
- golden vectors decoded by Pillow's BC7 decoder (tests/golden/make_bc7_vectors.py): every mode,
  every partition and every rotation / index selection, byte for byte, which pins the partition,
  anchor and weight tables entry by entry;
- the partition and anchor tables satisfy the format's invariants (texel 0 is in subset 0, every
  anchor texel lies in the subset it anchors, every partition uses all its subsets, no partition
  repeats);
- blocks packed here field by field for every mode decode to closed forms: equal endpoints give a
  constant colour (the endpoint widened by bit replication, whatever the indices), index 0 gives
  endpoint 0 and the top index endpoint 1, an anchor's index has one bit fewer, the rotation and
  index-selection fields of modes 4 and 5 exchange channels and index sets, and a block with no
  mode bit is transparent black."""
import numpy as np
import pytest

import oracle

# subsets, partition, rotation, selector, colour, alpha, endpoint p-bits, shared p-bits, index, index2
MODES = [(3, 4, 0, 0, 4, 0, 1, 0, 3, 0), (2, 6, 0, 0, 6, 0, 0, 1, 3, 0), (3, 6, 0, 0, 5, 0, 0, 0, 2, 0),
         (2, 6, 0, 0, 7, 0, 1, 0, 2, 0), (1, 0, 2, 1, 5, 6, 0, 0, 2, 3), (1, 0, 2, 0, 7, 8, 0, 0, 2, 2),
         (1, 0, 0, 0, 7, 7, 1, 0, 4, 0), (2, 6, 0, 0, 5, 5, 1, 0, 2, 0)]
WEIGHTS = {2: [0, 21, 43, 64], 3: [0, 9, 18, 27, 37, 46, 55, 64],
           4: [0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64]}


def widen(v, bits):
    v <<= 8 - bits
    return v | (v >> bits)


class Packer:
    def __init__(self):
        self.bits = []

    def put(self, value, n):
        self.bits += [(value >> k) & 1 for k in range(n)]

    def block(self):
        assert len(self.bits) == 128, len(self.bits)
        return bytes(sum(self.bits[8 * i + k] << k for k in range(8)) for i in range(16))


def pack(mode, partition=0, rotation=0, selector=0, ep=None, pbits=None, idx=None, idx2=None):
    """A mode-`mode` block from its fields: ep[subset][endpoint][channel] raw field values, pbits per
    subset (pairs, or one shared bit), idx / idx2 per texel (anchors written with one bit fewer)."""
    ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2 = MODES[mode]
    two, three, anchors = oracle.bc7_tables()
    anchor = {0}
    if ns == 2:
        anchor.add(int(anchors[0][partition]))
    if ns == 3:
        anchor |= {int(anchors[1][partition]), int(anchors[2][partition])}
    p = Packer()
    p.put(1 << mode, mode + 1)
    p.put(partition, pb)
    p.put(rotation, rb)
    p.put(selector, isb)
    for c in range(3):
        for s in range(ns):
            for e in range(2):
                p.put(ep[s][e][c], cb)
    for s in range(ns):
        for e in range(2):
            p.put(ep[s][e][3] if ab else 0, ab)
    for s in range(ns):
        if epb:
            p.put(pbits[s][0], 1)
            p.put(pbits[s][1], 1)
        elif spb:
            p.put(pbits[s], 1)
    for t in range(16):
        p.put(idx[t], ib - (1 if t in anchor else 0))
    if ib2:
        for t in range(16):
            p.put(idx2[t], ib2 - (1 if t == 0 else 0))
    return p.block()


def endpoint_value(mode, raw, pbit, alpha=False):
    ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2 = MODES[mode]
    bits = ab if alpha else cb
    if epb or spb:
        return widen((raw << 1) | pbit, bits + 1)
    return widen(raw, bits)


def decode1(block):
    return oracle.bc7_decode(block, 4, 4).reshape(16, 4).astype(int)


def test_partition_tables_satisfy_the_format_invariants():
    two, three, anchors = oracle.bc7_tables()
    assert (two[:, 0] == 0).all() and (three[:, 0] == 0).all()
    assert all(two[k][anchors[0][k]] == 1 for k in range(64))
    assert all(three[k][anchors[1][k]] == 1 and three[k][anchors[2][k]] == 2 for k in range(64))
    assert all(set(two[k]) == {0, 1} for k in range(64)) and all(set(three[k]) == {0, 1, 2} for k in range(64))
    assert len({tuple(r) for r in two}) == 64 and len({tuple(r) for r in three}) == 64


@pytest.mark.parametrize("mode", range(8))
def test_equal_endpoints_decode_to_the_widened_endpoint(mode):
    ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2 = MODES[mode]
    rng = np.random.default_rng(mode)
    for trial in range(20):
        part = int(rng.integers(0, 1 << pb)) if pb else 0
        raw = [[int(rng.integers(0, 1 << cb)) for _ in range(3)] + [int(rng.integers(0, 1 << ab)) if ab else 0]
               for _ in range(ns)]
        ep = [[raw[s], raw[s]] for s in range(ns)]
        pbits = [(p, p) for p in rng.integers(0, 2, ns)] if epb else [int(x) for x in rng.integers(0, 2, ns)]
        idx = [int(rng.integers(0, 1 << (ib - 1))) for _ in range(16)]
        idx2 = [int(rng.integers(0, 1 << (ib2 - 1))) for _ in range(16)] if ib2 else None
        got = decode1(pack(mode, part, 0, 0, ep, pbits, idx, idx2))
        two, three, _ = oracle.bc7_tables()
        for t in range(16):
            s = three[part][t] if ns == 3 else two[part][t] if ns == 2 else 0
            p = pbits[s][0] if epb else pbits[s] if spb else 0
            want = [endpoint_value(mode, raw[s][c], p) for c in range(3)]
            want.append(endpoint_value(mode, raw[s][3], p, alpha=True) if ab else 255)
            assert list(got[t]) == want, (mode, trial, t)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 6, 7])
def test_index_extremes_give_the_endpoints(mode):
    ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2 = MODES[mode]
    two, three, anchors = oracle.bc7_tables()
    part = 5 if pb else 0
    e0 = [[0, 0, 0, 0]] * ns
    e1 = [[(1 << cb) - 1] * 3 + [(1 << ab) - 1 if ab else 0]] * ns
    ep = [[e0[s], e1[s]] for s in range(ns)]
    pbits = [(0, 1)] * ns if epb else [0] * ns
    anchor = {0} | ({int(anchors[0][part])} if ns == 2 else set()) | (
        {int(anchors[1][part]), int(anchors[2][part])} if ns == 3 else set())
    top = (1 << ib) - 1
    idx = [(top >> 1) if t in anchor else top for t in range(16)]
    got = decode1(pack(mode, part, 0, 0, ep, pbits, idx))
    lo = endpoint_value(mode, 0, 0)
    hi = endpoint_value(mode, (1 << cb) - 1, 1 if epb else 0)
    for t in range(16):
        w = WEIGHTS[ib][idx[t]]
        want = ((64 - w) * lo + w * hi + 32) >> 6
        assert list(got[t][:3]) == [want] * 3, (mode, t)
        assert want == (hi if t not in anchor else want)
    # all indices 0: endpoint 0 everywhere
    got0 = decode1(pack(mode, part, 0, 0, ep, pbits, [0] * 16))
    assert (got0[:, :3] == lo).all()


@pytest.mark.parametrize("mode", [4, 5])
def test_rotation_and_index_selection(mode):
    ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2 = MODES[mode]
    rng = np.random.default_rng(40 + mode)
    for rotation in range(4):
        for selector in range(2 if isb else 1):
            c0 = [int(x) for x in rng.integers(0, 1 << cb, 3)] + [int(rng.integers(0, 1 << ab))]
            c1 = [int(x) for x in rng.integers(0, 1 << cb, 3)] + [int(rng.integers(0, 1 << ab))]
            idx = [int(rng.integers(0, 1 << (ib - (t == 0)))) for t in range(16)]
            idx2 = [int(rng.integers(0, 1 << (ib2 - (t == 0)))) for t in range(16)]
            got = decode1(pack(mode, 0, rotation, selector, [[c0, c1]], None, idx, idx2))
            for t in range(16):
                ci, cbits, ai, abits = idx[t], ib, idx2[t], ib2
                if selector:
                    ci, cbits, ai, abits = idx2[t], ib2, idx[t], ib
                px = []
                for c in range(4):
                    a, b = widen(c0[c], ab if c == 3 else cb), widen(c1[c], ab if c == 3 else cb)
                    w = WEIGHTS[abits][ai] if c == 3 else WEIGHTS[cbits][ci]
                    px.append(((64 - w) * a + w * b + 32) >> 6)
                if rotation:
                    px[rotation - 1], px[3] = px[3], px[rotation - 1]
                assert list(got[t]) == px, (mode, rotation, selector, t)


def test_partitions_select_the_subsets():
    two, three, _ = oracle.bc7_tables()
    for part in (0, 13, 37, 63):
        ep = [[[0, 0, 0, 0], [0, 0, 0, 0]], [[63, 63, 63, 0], [63, 63, 63, 0]]]
        got = decode1(pack(1, part, ep=ep, pbits=[1, 1], idx=[0] * 16))
        assert list((got[:, 0] == 255).astype(int)) == list(two[part])
        ep3 = [[[0] * 4] * 2, [[15, 0, 0, 0]] * 2, [[0, 15, 0, 0]] * 2]
        got = decode1(pack(0, part % 16, ep=ep3, pbits=[(1, 1)] * 3, idx=[0] * 16))
        want = [2 if g[1] == 255 else 1 if g[0] == 255 else 0 for g in got]
        assert want == list(three[part % 16])


def test_reserved_mode_and_image_edges():
    assert (decode1(bytes(16)) == 0).all()
    rng = np.random.default_rng(9)
    blocks = bytes(rng.integers(0, 256, 16 * 6, dtype=np.uint8))
    full = oracle.bc7_decode(blocks, 12, 8)
    crop = oracle.bc7_decode(blocks, 10, 7)
    assert np.array_equal(crop, full[:7, :10])


def test_every_mode_and_partition_matches_pillow():
    """Golden vectors from an independent decoder: 592 blocks covering every mode, all 16 / 64
    partitions of the two- and three-subset modes and every rotation / index selection of modes 4
    and 5, decoded by Pillow's BC7 decoder (tests/golden/make_bc7_vectors.py).  The oracle must
    reproduce every texel byte for byte, which pins its partition, anchor and weight tables."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "bc7_pillow.npz"))
    blocks, want = d["blocks"], d["rgba"]
    h, w = want.shape[:2]
    got = oracle.bc7_decode(blocks.tobytes(), w, h)
    bad = np.argwhere(np.any(got != want, axis=2))
    assert len(bad) == 0, f"{len(bad)} texels differ, first at {bad[:4].tolist()}"
    # the vectors reach every partition of every partitioned mode
    v = [int.from_bytes(b.tobytes(), "little") for b in blocks.reshape(-1, 16)]
    modes = [(x & -x).bit_length() - 1 for x in v]
    parts = {(m, (x >> (m + 1)) & (15 if m == 0 else 63)) for x, m in zip(v, modes) if m in (0, 1, 2, 3, 7)}
    assert len(parts) == 16 + 4 * 64
