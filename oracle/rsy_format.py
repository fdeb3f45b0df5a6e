"""CPU restatement of the .rsy asset format -- test infrastructure: the oracle the library's reader
and writer (rsy_asset_*) are checked against; nothing in the product imports it.

It follows the reference's definitions field by field with Python's struct module:
  - record layouts: file_header, material, sampler, surface, position (src/Asset/Asset.h:13-24,
    26-46, 48-54, 56-63, 92-99) with the C++ compiler's natural alignment (the material's two
    uint8_t fields are each followed by three padding bytes: 68 bytes in all), size_t as u64,
    little endian;
  - the write order of asset::write (src/Asset/Asset.cpp:39-403): header, six counts, materials,
    samplers, then per scene / node / image / mesh its size array and its data;
  - the checks of asset::read (Asset.cpp:405-848): magic, version, endianness, node and image size
    fields of 1, and a short read anywhere fails.

An asset is a dict: coordinate_system (16 floats), root_scene, materials (19-tuples in field
order, pads excluded), samplers (4-tuples), scenes (lists of node indices), nodes (dicts with
world_translate, world_scale, world_yaw, transform, mesh_id, children, name), images ((type, name)
pairs) and meshes ((positions as 16-float tuples, indices, surfaces as 9-tuples)).
"""
import struct

MAGIC = 0x52535946  # "RSYF", Asset.h:10
VERSION = 1         # Asset.h:11
LITTLE_ENDIAN = 1   # Asset.cpp:53

HEADER = struct.Struct("<3I16fI")        # magic, version, endianness, coordinate system, root scene
MATERIAL = struct.Struct("<B3x4f2fB3xf8I")  # double_sided, base colour, metallic, roughness, alpha mode, cutoff, 8 indices
SAMPLER = struct.Struct("<4H")
SURFACE = struct.Struct("<3I6f")
POSITION = struct.Struct("<16f")
U64 = struct.Struct("<Q")
assert (HEADER.size, MATERIAL.size, SAMPLER.size, SURFACE.size, POSITION.size) == (80, 68, 8, 36, 64)


class FormatError(ValueError):
    """The file is short, or a header or size field does not hold what asset::read requires."""


def encode(a):
    out = [HEADER.pack(MAGIC, VERSION, LITTLE_ENDIAN, *a["coordinate_system"], a["root_scene"])]
    out.append(struct.pack("<6Q", len(a["materials"]), len(a["samplers"]), len(a["scenes"]), len(a["nodes"]),
                           len(a["images"]), len(a["meshes"])))
    out += [MATERIAL.pack(*m) for m in a["materials"]]
    out += [SAMPLER.pack(*s) for s in a["samplers"]]
    for s in a["scenes"]:
        out.append(U64.pack(len(s)))
        out.append(struct.pack(f"<{len(s)}I", *s))
    for n in a["nodes"]:
        out.append(struct.pack("<7Q", 1, 1, 1, 1, 1, len(n["children"]), len(n["name"])))
        out.append(struct.pack("<3f", *n["world_translate"]))
        out.append(struct.pack("<ff", n["world_scale"], n["world_yaw"]))
        out.append(struct.pack("<16f", *n["transform"]))
        out.append(struct.pack("<I", n["mesh_id"]))
        out.append(struct.pack(f"<{len(n['children'])}I", *n["children"]))
        out.append(bytes(n["name"]))
    for image_type, name in a["images"]:
        out.append(struct.pack("<2Q", 1, len(name)))
        out.append(struct.pack("<I", image_type))
        out.append(bytes(name))
    for positions, indices, surfaces in a["meshes"]:
        out.append(struct.pack("<3Q", len(positions), len(indices), len(surfaces)))
        out += [POSITION.pack(*p) for p in positions]
        out.append(struct.pack(f"<{len(indices)}I", *indices))
        out += [SURFACE.pack(*s) for s in surfaces]
    return b"".join(out)


class _Cursor:
    def __init__(self, b):
        self.b, self.i = b, 0

    def take(self, n):
        if n < 0 or self.i + n > len(self.b):
            raise FormatError(f"short read: {n} bytes at offset {self.i} of {len(self.b)}")
        v = self.b[self.i:self.i + n]
        self.i += n
        return v

    def unpack(self, st):
        return st.unpack(self.take(st.size))

    def u64(self):
        return U64.unpack(self.take(8))[0]

    def u32s(self, n):
        return list(struct.unpack(f"<{n}I", self.take(4 * n)))


def decode(b):
    c = _Cursor(bytes(b))
    h = c.unpack(HEADER)
    if h[0] != MAGIC or h[1] != VERSION or h[2] != LITTLE_ENDIAN:
        raise FormatError("magic, version or endianness mismatch")
    a = {"coordinate_system": list(h[3:19]), "root_scene": h[19]}
    nm, ns, nsc, nn, ni, nme = struct.unpack("<6Q", c.take(48))
    a["materials"] = [c.unpack(MATERIAL) for _ in range(nm)]
    a["samplers"] = [c.unpack(SAMPLER) for _ in range(ns)]
    a["scenes"] = [c.u32s(c.u64()) for _ in range(nsc)]
    a["nodes"] = []
    for _ in range(nn):
        sizes = struct.unpack("<7Q", c.take(56))
        if sizes[:5] != (1, 1, 1, 1, 1):
            raise FormatError("node size fields other than 1")
        node = {"world_translate": list(struct.unpack("<3f", c.take(12)))}
        node["world_scale"], node["world_yaw"] = struct.unpack("<ff", c.take(8))
        node["transform"] = list(struct.unpack("<16f", c.take(64)))
        node["mesh_id"] = struct.unpack("<I", c.take(4))[0]
        node["children"] = c.u32s(sizes[5])
        node["name"] = c.take(sizes[6])
        a["nodes"].append(node)
    a["images"] = []
    for _ in range(ni):
        n_types, n_chars = struct.unpack("<2Q", c.take(16))
        if n_types != 1:
            raise FormatError("image type count other than 1")
        image_type = struct.unpack("<I", c.take(4))[0]
        a["images"].append((image_type, c.take(n_chars)))
    a["meshes"] = []
    for _ in range(nme):
        np_, ni_, ns_ = struct.unpack("<3Q", c.take(24))
        positions = [c.unpack(POSITION) for _ in range(np_)]
        indices = c.u32s(ni_)
        surfaces = [c.unpack(SURFACE) for _ in range(ns_)]
        a["meshes"].append((positions, indices, surfaces))
    return a


def boundaries(a):
    """Byte offsets at which asset::read issues a read (the start of every field group), in order,
    plus the file length: truncating a file at any offset below the length must fail."""
    offs = [0, HEADER.size]
    o = HEADER.size + 48
    offs.append(o)
    o += MATERIAL.size * len(a["materials"])
    offs.append(o)
    o += SAMPLER.size * len(a["samplers"])
    offs.append(o)
    for s in a["scenes"]:
        o += 8
        offs.append(o)
        o += 4 * len(s)
        offs.append(o)
    for n in a["nodes"]:
        for size in (56, 12, 4, 4, 64, 4, 4 * len(n["children"]), len(n["name"])):
            o += size
            offs.append(o)
    for image_type, name in a["images"]:
        for size in (16, 4, len(name)):
            o += size
            offs.append(o)
    for positions, indices, surfaces in a["meshes"]:
        for size in (24, 64 * len(positions), 4 * len(indices), 36 * len(surfaces)):
            o += size
            offs.append(o)
    return sorted(set(offs))
