"""Writes bc7_pillow.npz: BC7 blocks of every mode, every partition of the two- and three-subset
modes and every rotation / index-selection of modes 4 and 5, with the RGBA8 texels Pillow's own
BC7 decoder (PIL.DdsImagePlugin, an implementation independent of this repository's) gives them.

Run from the repository root: python tests/golden/make_bc7_vectors.py
The blocks' fields are random apart from the mode, partition, rotation and index-selection bits,
which are set so that every value of each is covered."""
import io
import os
import struct

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
DXGI_FORMAT_BC7_UNORM = 98


def block(rng, mode, extra_bits, extra_value):
    """Random 128-bit block with mode `mode` (bit `mode` set, those below clear) followed by
    `extra_bits` bits holding `extra_value` (partition, or rotation + index selection)."""
    v = int.from_bytes(rng.integers(0, 256, 16, dtype=np.uint8).tobytes(), "little")
    head = mode + 1 + extra_bits
    v &= ~((1 << head) - 1)
    v |= 1 << mode
    v |= extra_value << (mode + 1)
    return v.to_bytes(16, "little")


def dds_bc7(blocks, bw, bh):
    """A DDS file (DX10 header) holding bw x bh BC7 blocks."""
    w, h = 4 * bw, 4 * bh
    hdr = struct.pack("<4sIIIIIII44x", b"DDS ", 124, 0x1 | 0x2 | 0x4 | 0x1000 | 0x80000, h, w, len(blocks), 0, 1)
    pf = struct.pack("<II4sIIIII", 32, 0x4, b"DX10", 0, 0, 0, 0, 0)
    caps = struct.pack("<IIII4x", 0x1000, 0, 0, 0)
    dx10 = struct.pack("<IIIII", DXGI_FORMAT_BC7_UNORM, 3, 0, 1, 0)
    return hdr + pf + caps + dx10 + blocks


def main():
    rng = np.random.default_rng(20240611)
    # (mode, extra bits): partition bits for modes 0-3 and 7, rotation (+ index selection) for 4 and 5
    extras = {0: 4, 1: 6, 2: 6, 3: 6, 4: 3, 5: 2, 6: 0, 7: 6}
    out = []
    for mode, bits in extras.items():
        for value in range(1 << bits):
            for _ in range(3 if bits <= 3 else 2):
                out.append(block(rng, mode, bits, value))
    while len(out) % 16:  # a whole number of 16-block rows
        out.append(block(rng, 6, 0, 0))
    raw = b"".join(out)
    bw, bh = 16, len(out) // 16
    im = Image.open(io.BytesIO(dds_bc7(raw, bw, bh)))
    im.load()
    rgba = np.asarray(im.convert("RGBA"), np.uint8)
    assert rgba.shape == (4 * bh, 4 * bw, 4)
    blocks = np.frombuffer(raw, np.uint8).reshape(bh, bw, 16)
    np.savez_compressed(os.path.join(HERE, "bc7_pillow.npz"), blocks=blocks, rgba=rgba,
                        pillow=np.array(Image.__version__))
    print(f"{len(out)} blocks, {rgba.shape[1]}x{rgba.shape[0]} texels, Pillow {Image.__version__}")


if __name__ == "__main__":
    main()
