// .rsy asset files: reader and writer of the Packager's binary format (src/Asset/Asset.h:9-149,
// src/Asset/Asset.cpp:6-848), and the packing of a read asset into the frame path's geometry and
// material inputs (set_asset, src/Engine/Graphics.cpp:3525-4689).  Host code only.
//
// The file is the reference's structs written field group by field group with fwrite: little
// endian, natural alignment, size_t counts as u64.  Decoding checks every count against the bytes
// that remain before it allocates anything, so a truncated or corrupt file fails with
// RSY_READ_FAILED rather than a large allocation.
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "rosy_amd.h"

namespace {

constexpr uint32_t kMagic = 0x52535946u;  // "RSYF" (Asset.h:10)
constexpr uint32_t kVersion = 1u;         // Asset.h:11
constexpr uint32_t kLittleEndian = 1u;    // Asset.cpp:53, 461

// rosy_asset::file_header (Asset.h:13-24)
struct FileHeader {
    uint32_t magic;
    uint32_t version;
    uint32_t endianness;
    float coordinate_system[16];
    uint32_t root_scene;
};
static_assert(sizeof(FileHeader) == 80, "file header is 80 bytes");

// Byte sink of the encoder.
class Writer {
public:
    void put(const void* p, size_t n) {
        const auto* b = static_cast<const uint8_t*>(p);
        bytes.insert(bytes.end(), b, b + n);
    }
    void put_u64(uint64_t v) { put(&v, 8); }
    std::vector<uint8_t> bytes;
};

// Byte source of the decoder: every take is bounds-checked.
class Reader {
public:
    Reader(const uint8_t* p, uint64_t n) : p_(p), n_(n) {}
    bool take(void* dst, uint64_t n) {
        if (n > n_ - pos_) return false;
        if (n) std::memcpy(dst, p_ + pos_, size_t(n));
        pos_ += n;
        return true;
    }
    bool take_u64(uint64_t& v) { return take(&v, 8); }
    // room for `count` elements of `elem` bytes in what is left
    bool fits(uint64_t count, uint64_t elem) const { return elem == 0 || count <= (n_ - pos_) / elem; }

private:
    const uint8_t* p_;
    uint64_t n_;
    uint64_t pos_ = 0;
};

struct NodeData {
    float world_translate[3];
    float world_scale;
    float world_yaw;
    float transform[16];
    uint32_t mesh_id;
    std::vector<uint32_t> children;
    std::string name;
};

struct ImageData {
    uint32_t image_type;
    std::string name;
};

struct MeshData {
    std::vector<rsy_position> positions;
    std::vector<uint32_t> indices;
    std::vector<rsy_asset_surface> surfaces;
};

bool encode(const rsy_asset_desc& d, Writer& w) {
    // header (Asset.cpp:52-76)
    FileHeader h{};
    h.magic = kMagic;
    h.version = kVersion;
    h.endianness = kLittleEndian;
    std::memcpy(h.coordinate_system, d.coordinate_system, sizeof(h.coordinate_system));
    h.root_scene = d.root_scene;
    w.put(&h, sizeof(h));
    // six counts (Asset.cpp:80-107)
    const uint64_t counts[6] = {d.n_materials, d.n_samplers, d.n_scenes, d.n_nodes, d.n_images, d.n_meshes};
    w.put(counts, sizeof(counts));
    if ((d.n_materials && !d.materials) || (d.n_samplers && !d.samplers) || (d.n_scenes && !d.scenes) ||
        (d.n_nodes && !d.nodes) || (d.n_images && !d.images) || (d.n_meshes && !d.meshes))
        return false;
    w.put(d.materials, size_t(d.n_materials) * sizeof(rsy_asset_material));  // Asset.cpp:111-121
    w.put(d.samplers, size_t(d.n_samplers) * sizeof(rsy_sampler));           // Asset.cpp:124-134
    for (uint64_t i = 0; i < d.n_scenes; ++i) {                                // Asset.cpp:137-167
        const rsy_asset_scene& s = d.scenes[i];
        if (s.n_nodes && !s.nodes) return false;
        w.put_u64(s.n_nodes);
        w.put(s.nodes, size_t(s.n_nodes) * 4);
    }
    for (uint64_t i = 0; i < d.n_nodes; ++i) {  // Asset.cpp:170-270
        const rsy_asset_node& n = d.nodes[i];
        if ((n.n_child_nodes && !n.child_nodes) || (n.name_len && !n.name)) return false;
        const uint64_t sizes[7] = {1, 1, 1, 1, 1, n.n_child_nodes, n.name_len};
        w.put(sizes, sizeof(sizes));
        w.put(n.world_translate, 12);
        w.put(&n.world_scale, 4);
        w.put(&n.world_yaw, 4);
        w.put(n.transform, 64);
        w.put(&n.mesh_id, 4);
        w.put(n.child_nodes, size_t(n.n_child_nodes) * 4);
        w.put(n.name, size_t(n.name_len));
    }
    for (uint64_t i = 0; i < d.n_images; ++i) {  // Asset.cpp:273-308
        const rsy_asset_image& im = d.images[i];
        if (im.name_len && !im.name) return false;
        const uint64_t sizes[2] = {1, im.name_len};
        w.put(sizes, sizeof(sizes));
        w.put(&im.image_type, 4);
        w.put(im.name, size_t(im.name_len));
    }
    for (uint64_t i = 0; i < d.n_meshes; ++i) {  // Asset.cpp:311-361
        const rsy_asset_mesh& m = d.meshes[i];
        if ((m.n_positions && !m.positions) || (m.n_indices && !m.indices) || (m.n_surfaces && !m.surfaces))
            return false;
        const uint64_t sizes[3] = {m.n_positions, m.n_indices, m.n_surfaces};
        w.put(sizes, sizeof(sizes));
        w.put(m.positions, size_t(m.n_positions) * sizeof(rsy_position));
        w.put(m.indices, size_t(m.n_indices) * 4);
        w.put(m.surfaces, size_t(m.n_surfaces) * sizeof(rsy_asset_surface));
    }
    return true;
}

}  // namespace

struct rsy_asset {
    float coordinate_system[16];
    uint32_t root_scene = 0;
    std::vector<rsy_asset_material> materials;
    std::vector<rsy_sampler> samplers;
    std::vector<std::vector<uint32_t>> scenes;
    std::vector<NodeData> nodes;
    std::vector<ImageData> images;
    std::vector<MeshData> meshes;
    // views handed out by rsy_asset_describe
    std::vector<rsy_asset_scene> d_scenes;
    std::vector<rsy_asset_node> d_nodes;
    std::vector<rsy_asset_image> d_images;
    std::vector<rsy_asset_mesh> d_meshes;
};

namespace {

// asset::read (Asset.cpp:405-848) over an in-memory file.
rsy_result decode(const uint8_t* data, uint64_t size, rsy_asset& a) {
    Reader r(data, size);
    FileHeader h{};
    if (!r.take(&h, sizeof(h))) return RSY_READ_FAILED;
    if (h.magic != kMagic || h.version != kVersion || h.endianness != kLittleEndian) return RSY_READ_FAILED;
    std::memcpy(a.coordinate_system, h.coordinate_system, sizeof(a.coordinate_system));
    a.root_scene = h.root_scene;
    uint64_t counts[6];
    if (!r.take(counts, sizeof(counts))) return RSY_READ_FAILED;
    const uint64_t n_materials = counts[0], n_samplers = counts[1], n_scenes = counts[2];
    const uint64_t n_nodes = counts[3], n_images = counts[4], n_meshes = counts[5];
    // every scene, node, image and mesh takes at least its size fields
    if (!r.fits(n_materials, sizeof(rsy_asset_material))) return RSY_READ_FAILED;
    a.materials.resize(size_t(n_materials));
    if (!r.take(a.materials.data(), n_materials * sizeof(rsy_asset_material))) return RSY_READ_FAILED;
    if (!r.fits(n_samplers, sizeof(rsy_sampler))) return RSY_READ_FAILED;
    a.samplers.resize(size_t(n_samplers));
    if (!r.take(a.samplers.data(), n_samplers * sizeof(rsy_sampler))) return RSY_READ_FAILED;
    if (!r.fits(n_scenes, 8)) return RSY_READ_FAILED;
    a.scenes.resize(size_t(n_scenes));
    for (auto& s : a.scenes) {
        uint64_t n = 0;
        if (!r.take_u64(n) || !r.fits(n, 4)) return RSY_READ_FAILED;
        s.resize(size_t(n));
        if (!r.take(s.data(), n * 4)) return RSY_READ_FAILED;
    }
    if (!r.fits(n_nodes, 56)) return RSY_READ_FAILED;
    a.nodes.resize(size_t(n_nodes));
    for (auto& nd : a.nodes) {
        uint64_t sizes[7];
        if (!r.take(sizes, sizeof(sizes))) return RSY_READ_FAILED;
        // translate, scale, yaw, transform and mesh id are single values (asserted at Asset.cpp:587-591)
        for (int k = 0; k < 5; ++k)
            if (sizes[k] != 1) return RSY_READ_FAILED;
        if (!r.take(nd.world_translate, 12) || !r.take(&nd.world_scale, 4) || !r.take(&nd.world_yaw, 4) ||
            !r.take(nd.transform, 64) || !r.take(&nd.mesh_id, 4))
            return RSY_READ_FAILED;
        if (!r.fits(sizes[5], 4)) return RSY_READ_FAILED;
        nd.children.resize(size_t(sizes[5]));
        if (!r.take(nd.children.data(), sizes[5] * 4)) return RSY_READ_FAILED;
        if (!r.fits(sizes[6], 1)) return RSY_READ_FAILED;
        nd.name.resize(size_t(sizes[6]));
        if (!r.take(&nd.name[0], sizes[6])) return RSY_READ_FAILED;
    }
    if (!r.fits(n_images, 16)) return RSY_READ_FAILED;
    a.images.resize(size_t(n_images));
    for (auto& im : a.images) {
        uint64_t sizes[2];
        if (!r.take(sizes, sizeof(sizes)) || sizes[0] != 1) return RSY_READ_FAILED;  // Asset.cpp:698
        if (!r.take(&im.image_type, 4) || !r.fits(sizes[1], 1)) return RSY_READ_FAILED;
        im.name.resize(size_t(sizes[1]));
        if (!r.take(&im.name[0], sizes[1])) return RSY_READ_FAILED;
    }
    if (!r.fits(n_meshes, 24)) return RSY_READ_FAILED;
    a.meshes.resize(size_t(n_meshes));
    for (auto& m : a.meshes) {
        uint64_t sizes[3];
        if (!r.take(sizes, sizeof(sizes))) return RSY_READ_FAILED;
        if (!r.fits(sizes[0], sizeof(rsy_position))) return RSY_READ_FAILED;
        m.positions.resize(size_t(sizes[0]));
        if (!r.take(m.positions.data(), sizes[0] * sizeof(rsy_position))) return RSY_READ_FAILED;
        if (!r.fits(sizes[1], 4)) return RSY_READ_FAILED;
        m.indices.resize(size_t(sizes[1]));
        if (!r.take(m.indices.data(), sizes[1] * 4)) return RSY_READ_FAILED;
        if (!r.fits(sizes[2], sizeof(rsy_asset_surface))) return RSY_READ_FAILED;
        m.surfaces.resize(size_t(sizes[2]));
        if (!r.take(m.surfaces.data(), sizes[2] * sizeof(rsy_asset_surface))) return RSY_READ_FAILED;
    }
    return RSY_OK;
}

bool read_file(const char* path, std::vector<uint8_t>& bytes) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long n = ok ? std::ftell(f) : -1;
    ok = ok && n >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        bytes.resize(size_t(n));
        ok = n == 0 || std::fread(bytes.data(), 1, size_t(n), f) == size_t(n);
    }
    std::fclose(f);
    return ok;
}

}  // namespace

extern "C" {

rsy_result rsy_asset_encode(const rsy_asset_desc* desc, void* out, uint64_t capacity, uint64_t* size) {
    if (!desc || !size) return RSY_INVALID_ARGUMENT;
    try {
        Writer w;
        if (!encode(*desc, w)) return RSY_INVALID_ARGUMENT;
        *size = w.bytes.size();
        if (!out) return RSY_OK;
        if (capacity < w.bytes.size()) return RSY_INVALID_ARGUMENT;
        std::memcpy(out, w.bytes.data(), w.bytes.size());
        return RSY_OK;
    } catch (const std::bad_alloc&) {
        return RSY_ALLOCATION_FAILURE;
    }
}

rsy_result rsy_asset_write(const rsy_asset_desc* desc, const char* path) {
    if (!desc || !path) return RSY_INVALID_ARGUMENT;
    try {
        Writer w;
        if (!encode(*desc, w)) return RSY_INVALID_ARGUMENT;
        FILE* f = std::fopen(path, "wb");
        if (!f) return RSY_OPEN_FAILED;
        const bool ok = w.bytes.empty() || std::fwrite(w.bytes.data(), 1, w.bytes.size(), f) == w.bytes.size();
        const bool closed = std::fclose(f) == 0;
        return ok && closed ? RSY_OK : RSY_WRITE_FAILED;
    } catch (const std::bad_alloc&) {
        return RSY_ALLOCATION_FAILURE;
    }
}

rsy_result rsy_asset_decode(const void* data, uint64_t size, rsy_asset** out_asset) {
    if (!out_asset || (!data && size)) return RSY_INVALID_ARGUMENT;
    *out_asset = nullptr;
    rsy_asset* a = new (std::nothrow) rsy_asset();
    if (!a) return RSY_ALLOCATION_FAILURE;
    rsy_result r;
    try {
        r = decode(static_cast<const uint8_t*>(data), size, *a);
    } catch (const std::bad_alloc&) {
        r = RSY_ALLOCATION_FAILURE;
    }
    if (r != RSY_OK) {
        delete a;
        return r;
    }
    *out_asset = a;
    return RSY_OK;
}

rsy_result rsy_asset_read(const char* path, rsy_asset** out_asset) {
    if (!path || !out_asset) return RSY_INVALID_ARGUMENT;
    *out_asset = nullptr;
    std::vector<uint8_t> bytes;
    try {
        if (!read_file(path, bytes)) return RSY_OPEN_FAILED;
    } catch (const std::bad_alloc&) {
        return RSY_ALLOCATION_FAILURE;
    }
    return rsy_asset_decode(bytes.data(), bytes.size(), out_asset);
}

rsy_result rsy_asset_describe(const rsy_asset* asset, rsy_asset_desc* out) {
    if (!asset || !out) return RSY_INVALID_ARGUMENT;
    rsy_asset* a = const_cast<rsy_asset*>(asset);
    try {
        a->d_scenes.resize(a->scenes.size());
        for (size_t i = 0; i < a->scenes.size(); ++i) a->d_scenes[i] = {a->scenes[i].data(), a->scenes[i].size()};
        a->d_nodes.resize(a->nodes.size());
        for (size_t i = 0; i < a->nodes.size(); ++i) {
            const NodeData& n = a->nodes[i];
            rsy_asset_node& d = a->d_nodes[i];
            std::memcpy(d.world_translate, n.world_translate, sizeof(d.world_translate));
            d.world_scale = n.world_scale;
            d.world_yaw = n.world_yaw;
            std::memcpy(d.transform, n.transform, sizeof(d.transform));
            d.mesh_id = n.mesh_id;
            d.child_nodes = n.children.data();
            d.n_child_nodes = n.children.size();
            d.name = n.name.data();
            d.name_len = n.name.size();
        }
        a->d_images.resize(a->images.size());
        for (size_t i = 0; i < a->images.size(); ++i)
            a->d_images[i] = {a->images[i].image_type, a->images[i].name.data(), a->images[i].name.size()};
        a->d_meshes.resize(a->meshes.size());
        for (size_t i = 0; i < a->meshes.size(); ++i) {
            const MeshData& m = a->meshes[i];
            a->d_meshes[i] = {m.positions.data(), m.positions.size(), m.indices.data(), m.indices.size(),
                              m.surfaces.data(), m.surfaces.size()};
        }
    } catch (const std::bad_alloc&) {
        return RSY_ALLOCATION_FAILURE;
    }
    std::memcpy(out->coordinate_system, a->coordinate_system, sizeof(out->coordinate_system));
    out->root_scene = a->root_scene;
    out->materials = a->materials.data();
    out->n_materials = a->materials.size();
    out->samplers = a->samplers.data();
    out->n_samplers = a->samplers.size();
    out->scenes = a->d_scenes.data();
    out->n_scenes = a->d_scenes.size();
    out->nodes = a->d_nodes.data();
    out->n_nodes = a->d_nodes.size();
    out->images = a->d_images.data();
    out->n_images = a->d_images.size();
    out->meshes = a->d_meshes.data();
    out->n_meshes = a->d_meshes.size();
    return RSY_OK;
}

void rsy_asset_destroy(rsy_asset* asset) { delete asset; }

rsy_result rsy_asset_geometry_size(const rsy_asset* asset, uint64_t* n_vertices, uint64_t* n_indices,
                                   uint32_t* n_meshes) {
    if (!asset || !n_vertices || !n_indices || !n_meshes) return RSY_INVALID_ARGUMENT;
    uint64_t nv = 0, ni = 0;
    for (const MeshData& m : asset->meshes) {
        nv += m.positions.size();
        ni += m.indices.size();
    }
    if (asset->meshes.size() > 0xffffffffull) return RSY_OVERFLOW;
    *n_vertices = nv;
    *n_indices = ni;
    *n_meshes = uint32_t(asset->meshes.size());
    return RSY_OK;
}

rsy_result rsy_asset_pack_geometry(const rsy_asset* asset, rsy_position* vertices, uint64_t n_vertices,
                                   uint32_t* indices, uint64_t n_indices, rsy_mesh_range* meshes, uint32_t n_meshes) {
    uint64_t nv = 0, ni = 0;
    uint32_t nm = 0;
    const rsy_result r = rsy_asset_geometry_size(asset, &nv, &ni, &nm);
    if (r != RSY_OK) return r;
    if (nv != n_vertices || ni != n_indices || nm != n_meshes) return RSY_INVALID_ARGUMENT;
    if ((nv && !vertices) || (ni && !indices) || (nm && !meshes)) return RSY_INVALID_ARGUMENT;
    // index offsets are u32 in the reference's gpu_mesh_buffers (Graphics.cpp:4313-4326)
    if (ni > 0xffffffffull) return RSY_OVERFLOW;
    uint64_t v = 0, i = 0;
    for (uint32_t k = 0; k < nm; ++k) {
        const MeshData& m = asset->meshes[k];
        if (m.indices.size() > 0xffffffffull || m.positions.size() > 0xffffffffull) return RSY_OVERFLOW;
        rsy_mesh_range& mr = meshes[k];
        mr.vertex_buffer_offset = v * sizeof(rsy_position);
        mr.index_offset = uint32_t(i);
        mr.num_indices = uint32_t(m.indices.size());
        mr.num_vertices = uint32_t(m.positions.size());
        mr.reserved = 0;
        if (!m.positions.empty()) std::memcpy(vertices + v, m.positions.data(), m.positions.size() * sizeof(rsy_position));
        if (!m.indices.empty()) std::memcpy(indices + i, m.indices.data(), m.indices.size() * 4);
        v += m.positions.size();
        i += m.indices.size();
    }
    return RSY_OK;
}

rsy_result rsy_asset_pack_materials(const rsy_asset* asset, rsy_gpu_material* materials, uint32_t n_materials) {
    if (!asset || (n_materials && !materials)) return RSY_INVALID_ARGUMENT;
    if (n_materials != asset->materials.size()) return RSY_INVALID_ARGUMENT;
    const uint64_t n_images = asset->images.size(), n_samplers = asset->samplers.size();
    // one map of the material: the image if it is in range, its sampler if that is in range too
    auto map = [&](uint32_t image, uint32_t sampler, uint32_t& out_image, uint32_t& out_sampler) {
        out_image = RSY_ABSENT;
        out_sampler = RSY_ABSENT;  // the default sampler (REPEAT, linear; Graphics.cpp:3421-3432)
        if (image < n_images) {
            out_image = image;
            if (sampler < n_samplers) out_sampler = sampler;
        }
    };
    for (uint32_t k = 0; k < n_materials; ++k) {
        const rsy_asset_material& m = asset->materials[k];
        rsy_gpu_material& g = materials[k];
        std::memcpy(g.color, m.base_color_factor, sizeof(g.color));
        g.metallic_factor = m.metallic_factor;
        g.roughness_factor = m.roughness_factor;
        g.alpha_cutoff = m.alpha_cutoff;
        g.alpha_mode = m.alpha_mode;
        map(m.color_image_index, m.color_sampler_index, g.color_sampled_image_index, g.color_sampler_index);
        map(m.normal_image_index, m.normal_sampler_index, g.normal_sampled_image_index, g.normal_sampler_index);
        map(m.metallic_image_index, m.metallic_sampler_index, g.metallic_sampled_image_index,
            g.metallic_sampler_index);
        map(m.mixmap_image_index, m.mixmap_sampler_index, g.mixmap_sampled_image_index, g.mixmap_sampler_index);
    }
    return RSY_OK;
}

}  // extern "C"
