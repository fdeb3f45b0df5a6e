#!/usr/bin/env python3
"""Frame benchmark of the raster-and-shade path on one MI355X: plain vs neural-material shading.

Workload (seeded, synthetic): 120 procedural meshes, 262,144 triangles, one 1920x1080 camera view
and a 2048^2 shadow map; 64 materials whose RGBA8 albedo, RG8 normal and RG8 metallic-roughness
maps are 1024^2 (value noise + stripes + blobs).  The neural material decoder (256^2 x 8-channel
FP16 latent grid per material + one 10-32-7 MLP) is trained on the GPU first (Adam, lr 1e-3 with
cosine decay, 20,000 steps of 65,536 samples), its PSNR measured per material, and every material
at or above 40 dB decoded; the rest keep their plain maps.

A step is one whole frame: shadow pass, main visibility pass and deferred Cook-Torrance + PCF
shading, replayed from a captured HIP graph with every input resident in HBM.  `value` is the
neural-material frame in device milliseconds (HIP events around the timed replays); the plain frame
is timed the same way beside it.

Also reported:
  roofline      the kernel that dominates the frame, the shading kernel of the neural form: its
                algorithmic bytes per launch (8 B visibility read and 16 B colour write per pixel,
                48 B per visible vertex, the unique latent cells (16 B) and map texels it reads, 4 B
                per unique shadow-map texel; the unique counts come from the oracle's footprint of
                the same frame) / its average duration measured live with HIP events on the
                context's stream, against 8 TB/s; `traffic` is its measured HBM bytes per launch from
                the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
                (profiles/k_shade_pmc_decoder_1080p.json);
  kernels       the same for the plain shading kernel, the tile raster (4 B per (triangle, bin)
                pair, 32 B per set-up triangle, 8 B per camera pixel and 4 B per shadow texel of the
                bins it writes) and the whole geometry + binning + raster stage (the survey's B_geo);
  decoder       MFMA FLOPs of the decoder per frame / the neural shading kernel's duration, against
                the 2.5 PFLOP/s dense FP16 peak (the shading kernel is not MFMA-bound);
  cpu_baseline  the CPU oracle (fixed-point raster + FP64 shading, std::thread over the host cores)
                on the same plain frame, repeated for a bounded time;
  max_abs_err   shaded colour of both forms vs the FP64 CPU oracle, with the parity tests' rule for
                shadow compares: the GPU's PCF factor may differ from the oracle's only at a compare
                tie (|reference - texel| < 1e-5), and the colour is compared with the oracle re-run
                with the GPU's factor (for the neural form with the same FP16 decoder parameters).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/frame @1080p,262k tris: plain vs neural-material shading @PSNR>=40dB"
HBM_PEAK_GBS = 8000.0
MFMA_FP16_PEAK_TFLOPS = 2500.0
DECODER_FLOPS_PER_PIXEL = 2 * (10 * 32 + 32 * 7)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def choose_tile(r, scene, frame, tile):
    """Bin size of the run: the autotuner's pick (tile 0) or the given size, fixed in the raster state."""
    if tile == 0:
        tile = r.autotune(frame, 10)
        log(f"autotuned bin size {tile}")
        return tile
    import ctypes as C

    from rosy_amd import abi
    rs = abi.RasterState()
    C.memmove(C.byref(rs), C.byref(scene.s.raster_state), C.sizeof(rs))
    rs.tile_size = tile
    r.update(scene.s.scene_data, rs)
    log(f"bin size {tile} (fixed)")
    return tile


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_shade_pmc_decoder_1080p.json")


def measured_traffic(kernel, tile):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE passes of this
    workload (each counter scaled by what it reads for a plain device copy of known size), or None
    when no measurement of this kernel at this bin size is committed."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if k is None or t.get("tile") != tile:
        return None
    return int(round(1024 * (k["fetch_kib_per_dispatch"] * t["fetch_scale"] +
                             k["write_kib_per_dispatch"] * t["write_scale"])))


def hbm_line(kernel, nbytes, ms, extra=None):
    """One kernel's HBM roofline entry: algorithmic bytes per launch over its measured duration."""
    gbs = nbytes / (ms * 1e-3) / 1e9
    d = {"kernel": kernel, "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(gbs / HBM_PEAK_GBS, 5), "algorithmic_bytes": int(nbytes), "avg_launch_ms": round(ms, 5)}
    if extra:
        d.update(extra)
    return d


def shade_bytes(fp, pixels):
    """Algorithmic bytes of one shading launch from the oracle's footprint: 8 B visibility read and
    16 B colour write per pixel, 48 B per visible vertex (normal, tangent, uv, colour), the unique
    map texels, 16 B per unique latent cell (8 FP16 channels), 4 B per unique shadow-map texel."""
    parts = {"pixels": 24 * pixels, "vertices": 48 * fp["visible_vertices"], "texels": fp["texel_bytes"],
             "latents": 16 * fp["latent_cells"], "shadow": 4 * fp["shadow_texels"]}
    return sum(parts.values()), parts


def list_vertices(scene):
    """Vertices the transform reads for the opaque list (every surface of these workloads)."""
    return sum(int(scene.s.meshes[d.mesh_index].num_vertices) for d in scene.surfaces())


def geo_bytes(scene, views, pairs):
    """The survey's B_geo over the frame's views: per view 12 V (positions) + 12 T (indices) +
    164 per object + 2 x 16 V (transformed vertices written and read) + 2 x 32 T (setup records
    written and read) + 8 B per visibility word or 4 B per shadow texel; plus 8 B per (triangle, bin)
    pair over the frame (list entry written and read)."""
    V, T, n_obj = list_vertices(scene), scene.n_triangles, scene.s.n_objects
    b = 0
    for w, h, shadow in views:
        b += 12 * V + 12 * T + 164 * n_obj + 32 * V + 64 * T + w * h * (4 if shadow else 8)
    return b + 8 * pairs


def raster_bytes(scene, views, rs, tile, pairs, setup):
    """The tile raster's algorithmic bytes: 4 B per (triangle, bin) pair, 32 B per set-up triangle,
    8 B per camera pixel and 4 B per shadow texel of the bins it writes (those with a triangle; empty
    bins are cleared by the scatter)."""
    import numpy as np

    import oracle
    b = 4 * pairs + 32 * setup
    for vp, w, h, shadow in views:
        counts, _ = oracle.tile_counts(scene, vp, w, h, shadow, rs, tile)
        ys, xs = np.nonzero(counts)
        px = sum(min(tile, w - x * tile) * min(tile, h - y * tile) for y, x in zip(ys.tolist(), xs.tolist()))
        b += px * (4 if shadow else 8)
    return b


def colour_error(col, sf, ref_fn):
    """Max abs colour error vs the FP64 oracle with the parity tests' PCF rule: the oracle's own
    PCF factor must equal the GPU's except at a compare tie (|ref - texel| < 1e-5); the colour is then
    compared with the oracle re-run with the GPU's factor.  Returns (error, tie pixels, pixels whose
    factor differs away from a tie)."""
    import numpy as np
    ref, rsf, margin = ref_fn(None)
    flip = np.abs(sf - rsf) > 2e-3
    ties = int(flip.sum())
    away = int((flip & ~(margin < 1e-5)).sum())
    ref2, _, _ = ref_fn(sf)
    return float(np.abs(col.astype(np.float64) - ref2).max()), ties, away


def stage_medians(r, frame, n=10):
    import numpy as np
    from rosy_amd import abi
    r.set_debug_flags(abi.DEBUG_STAGE_TIMING | abi.DEBUG_SHADOW_FACTOR)
    st = []
    for _ in range(n):
        r.render(frame)
        st.append(list(r.stats().stage_ms)[:5])
    return np.median(np.array(st), axis=0)


def present_timing(r, width, height, n=10):
    """Output stage of view 0 to a width x height B8G8R8A8_SRGB image: median device ms of the draw
    image conversion + blit (HIP events around the two kernels) and its algorithmic bytes (16 B
    colour read + 8 B draw image write + 8 B draw image read per source pixel, 4 B per destination
    pixel)."""
    import numpy as np
    from rosy_amd import abi
    r.set_debug_flags(abi.DEBUG_STAGE_TIMING)
    t = []
    for _ in range(n):
        r.present(0, width, height, readback=False)
        t.append(float(r.stats().stage_ms[5]))
    r.set_debug_flags(0)
    w, h = r.views[0]
    ms = float(np.median(t))
    nbytes = w * h * 32 + width * height * 4
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"extent": f"{w}x{h} -> {width}x{height}", "ms": round(ms, 5), "algorithmic_bytes": nbytes,
            "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def msaa_timing(r, scene, plain, neural, args, check):
    """Graph-replayed frames with 4 samples per camera pixel (bin size 32), both shading forms, and
    the stage medians of the neural one; with `check`, every sample plane against the oracle."""
    import ctypes as C

    import numpy as np
    from rosy_amd import abi
    rs = abi.RasterState()
    C.memmove(C.byref(rs), C.byref(scene.s.raster_state), C.sizeof(rs))
    rs.msaa_samples = 4
    rs.tile_size = 32
    r.update(scene.s.scene_data, rs)
    st = stage_medians(r, neural)
    exact = None
    if check:
        import oracle
        cam = scene.camera_views()[0]
        want = oracle.raster_view_msaa(scene, cam.view_projection[:], cam.width, cam.height, rs, oracle.threads())
        exact = bool(np.array_equal(r.visibility(0), want))
    r.set_debug_flags(0)
    out = {}
    for name, frame in (("plain", plain), ("neural", neural)):
        r.render_graph(frame, args.warmup)
        r.render_graph(frame, args.steps)
        out[name + "_ms_per_frame"] = round(float(r.stats().gpu_ms), 5)
    out["stage_ms_neural"] = {"geometry": round(float(st[0]), 5), "binning": round(float(st[1]), 5),
                              "raster": round(float(st[2]), 5), "shade": round(float(st[3]), 5)}
    out["sample_planes_bit_exact_vs_oracle"] = exact
    rs.msaa_samples = 0
    rs.tile_size = scene.s.raster_state.tile_size
    r.update(scene.s.scene_data, rs)
    return out


def copy_ceiling_gbs(mib=1024, reps=10):
    """A measured HBM ceiling on this box: device-to-device copies of `mib` MiB (read + write bytes per
    second, HIP events), the best of `reps`.  Through the HIP runtime the library itself uses."""
    import ctypes as C

    import rosy_amd
    rosy_amd.lib()  # the library's HIP runtime is loaded: use the same one
    hip = C.CDLL("libamdhip64.so.7")  # the soname the library links
    n = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(n)) or hip.hipMalloc(C.byref(dst), C.c_size_t(n)):
        return None
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemset(src, 1, C.c_size_t(n))
    best = 0.0
    for _ in range(reps):
        hip.hipEventRecord(e0, None)
        hip.hipMemcpyAsync(dst, src, C.c_size_t(n), 3, None)  # hipMemcpyDeviceToDevice
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        if ms.value > 0:
            best = max(best, 2.0 * n / (ms.value * 1e-3) / 1e9)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(src)
    hip.hipFree(dst)
    return round(best, 1)


def texture_timing(args):
    """BC7 images with full chains uploaded through rsy_set_materials (64 x 1024^2, random blocks of
    every mode: host-to-device copy plus GPU decode, wall clock), and the plain decoder-workload frame
    with every texture's chain built (trilinear sampling at the pixel's level of detail)."""
    import ctypes as C

    import numpy as np
    import rosy_amd
    from rosy_amd import abi
    rng = np.random.default_rng(11)
    size, n_tex = 1024, 64
    levels = int(np.log2(size)) + 1
    nblocks = sum(max(1, (size >> l) // 4) ** 2 for l in range(levels))
    blocks = rng.integers(0, 256, (nblocks, 16), dtype=np.uint8)
    blocks[:, 0] = np.left_shift(1, rng.integers(0, 8, nblocks)).astype(np.uint8) | (blocks[:, 0] & 0x80)
    data = blocks.tobytes()
    descs = (abi.TextureDesc * n_tex)()
    for i in range(n_tex):
        descs[i] = abi.TextureDesc(size, size, abi.TEX_BC7_SRGB, levels, C.cast(C.c_char_p(data), C.c_void_p))
    small = rosy_amd.Scene(0, seed=1)
    r = rosy_amd.Renderer(device=0)
    r.load(small)
    lib = rosy_amd.lib()
    lib.rsy_set_materials(r.ctx, None, 0, descs, n_tex, None, 0)  # warm-up: allocations
    t0 = time.perf_counter()
    rc = lib.rsy_set_materials(r.ctx, None, 0, descs, n_tex, None, 0)
    upload_ms = (time.perf_counter() - t0) * 1e3
    assert rc == abi.RSY_OK
    r.close()
    small.close()
    out = {"bc7_upload": {"textures": n_tex, "extent": size, "levels": levels, "mb": round(n_tex * len(data) / 1e6, 1),
                          "ms": round(upload_ms, 2), "gbs": round(n_tex * len(data) / (upload_ms * 1e-3) / 1e9, 2),
                          "note": "wall clock of rsy_set_materials: host-to-device copy and GPU decode"}}
    scene = rosy_amd.Scene(3, seed=args.seed, texture_mips=1)
    r = rosy_amd.Renderer(device=0)
    r.load(scene)
    plain = scene.frame_desc(abi.SHADE_PLAIN)
    st = stage_medians(r, plain)
    r.set_debug_flags(0)
    r.render_graph(plain, args.warmup)
    r.render_graph(plain, args.steps)
    out["mip_chains"] = {"plain_ms_per_frame": round(float(r.stats().gpu_ms), 5), "shade_ms": round(float(st[3]), 5),
                         "workload": "decoder workload scene, every 1024^2 map with its built chain (11 levels)"}
    r.close()
    scene.close()
    return out


def _srgb_lut():
    import numpy as np
    v = np.arange(256) / 255.0
    return np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)


def _map_channels(albedo, normal, mr, lut):
    """The decoder's 7 channels from a material's three stored maps (the plain sampler's values at
    texel centres): linear albedo rgb, normal-map r g (UNORM), metallic and roughness (sRGB decoded)."""
    import numpy as np
    return np.concatenate([lut[albedo[..., :3]], normal[..., :2] / 255.0, lut[mr[..., :2]]], axis=-1)


def _bilinear_repeat(img, out_w, out_h):
    """Bilinear REPEAT sampling of img (h, w, c) at the texel centres of an out_w x out_h grid."""
    import numpy as np
    h, w = img.shape[:2]

    def axis(n_out, n_in):
        x = (np.arange(n_out) + 0.5) / n_out * n_in - 0.5
        x0 = np.floor(x).astype(np.int64)
        return x0 % n_in, (x0 + 1) % n_in, (x - x0)

    xa, xb, wx = axis(out_w, w)
    ya, yb, wy = axis(out_h, h)
    wx = wx[None, :, None]
    top = img[ya][:, xa] * (1 - wx) + img[ya][:, xb] * wx
    bot = img[yb][:, xa] * (1 - wx) + img[yb][:, xb] * wx
    return top * (1 - wy[:, None, None]) + bot * wy[:, None, None]


def low_res_ablation(scene, frame_plain_colour, covered, args, level=2):
    """The method without the network: plain shading from box-filtered copies of the maps at the
    latent grid's resolution (level `level` of the packager-filter chain the library builds: 1024^2
    -> 256^2).  Reports the shading kernel's time with those maps, the frame PSNR against the
    full-resolution plain frame, and the texel-centre PSNR per material over the decoder's 7
    channels (the decoder's own metric)."""
    import ctypes as C

    import numpy as np
    import rosy_amd
    from rosy_amd import abi
    s = scene.s
    r = rosy_amd.Renderer(device=0)
    lib = rosy_amd.lib()
    try:
        r.load(scene)
        n_tex = s.n_textures
        descs = (abi.TextureDesc * n_tex)()
        for t in range(n_tex):
            d = s.textures[t]
            descs[t] = abi.TextureDesc(d.width, d.height, d.format, abi.MIPS_BUILD, d.pixels)
        assert lib.rsy_set_materials(r.ctx, s.materials, s.n_materials, descs, n_tex, s.samplers, s.n_samplers) == 0
        low, keep = [], []
        for t in range(n_tex):
            d = s.textures[t]
            ch = 2 if d.format in (abi.TEX_RG8_UNORM, abi.TEX_RG8_SRGB) else 4
            w, h = max(1, d.width >> level), max(1, d.height >> level)
            buf = np.empty(w * h * ch, np.uint8)
            assert lib.rsy_read_texture(r.ctx, t, level, buf.ctypes.data, buf.size) == 0
            low.append(buf.reshape(h, w, ch))
            keep.append(buf)
            descs[t] = abi.TextureDesc(w, h, d.format, 1, buf.ctypes.data)
        assert lib.rsy_set_materials(r.ctx, s.materials, s.n_materials, descs, n_tex, s.samplers, s.n_samplers) == 0
        plain = scene.frame_desc(abi.SHADE_PLAIN)
        st = stage_medians(r, plain)
        col = r.color(0)
        r.set_debug_flags(0)
        r.render_graph(plain, args.warmup)
        r.render_graph(plain, args.steps)
        frame_ms = float(r.stats().gpu_ms)
    finally:
        r.close()
    diff = col[covered][:, :3].astype(np.float64) - frame_plain_colour[covered][:, :3]
    frame_psnr = float(10 * np.log10(1.0 / np.mean(diff ** 2)))
    # texel-centre PSNR per material, the decoder's metric (rsy_neural_evaluate)
    lut = _srgb_lut()
    psnr = []
    for m in range(s.n_materials):
        md = s.materials[m]
        ia, inn, imr = md.color_sampled_image_index, md.normal_sampled_image_index, md.metallic_sampled_image_index
        d = s.textures[ia]
        full = [np.ctypeslib.as_array(C.cast(s.textures[i].pixels, C.POINTER(C.c_uint8)),
                                      (s.textures[i].height, s.textures[i].width,
                                       2 if s.textures[i].format in (abi.TEX_RG8_UNORM, abi.TEX_RG8_SRGB) else 4))
                for i in (ia, inn, imr)]
        want = _map_channels(*full, lut)
        got = _bilinear_repeat(_map_channels(low[ia], low[inn], low[imr], lut), d.width, d.height)
        psnr.append(float(10 * np.log10(1.0 / np.mean((got - want) ** 2))))
    psnr = np.array(psnr)
    return {"maps": f"{s.textures[0].width >> level}^2 box-filtered copies (the packager filter, level {level})",
            "shade_ms": round(float(st[3]), 5), "plain_ms_per_frame": round(frame_ms, 5),
            "frame_psnr_vs_full_res_plain_db": round(frame_psnr, 2),
            "texel_psnr_min": round(float(psnr.min()), 2), "texel_psnr_median": round(float(np.median(psnr)), 2),
            "materials_at_40db": int((psnr >= 40.0).sum())}


def decoder_study(args, detail, train_steps):
    """One decoder workload: the decoder scene with the default content (detail 0) or with fine detail
    (detail 1: content above a 256^2 grid's Nyquist rate).  Trains on the GPU, gates at 40 dB
    (materials below keep their plain maps), times both frame forms from graph replays, and compares
    with the low-resolution-map ablation."""
    import numpy as np
    import rosy_amd
    from rosy_amd import abi
    scene = rosy_amd.Scene(3, seed=args.seed, texture_detail=detail)
    r = rosy_amd.Renderer(device=0)
    try:
        r.load(scene)
        plain = scene.frame_desc(abi.SHADE_PLAIN)
        neural = scene.frame_desc(abi.SHADE_NEURAL)
        r.neural_init(decay_steps=train_steps)
        psnr0 = r.neural_evaluate()
        r.neural_train(train_steps)
        psnr = r.neural_evaluate()
        nst = r.neural_stats()
        st_plain = stage_medians(r, plain)
        col_plain = r.color(0)
        covered = r.visibility(0) != 0
        st_neural = stage_medians(r, neural)
        col_neural = r.color(0)
        r.set_debug_flags(0)
        ms = {}
        for name, f in (("plain", plain), ("neural", neural)):
            r.render_graph(f, args.warmup)
            r.render_graph(f, args.steps)
            ms[name] = float(r.stats().gpu_ms)
        diff = col_neural[covered][:, :3].astype(np.float64) - col_plain[covered][:, :3]
        out = {"content": "fine detail (up to 512 cycles per map, stripes up to 200 cycles)" if detail else
               "default (up to 32 cycles per map, stripes of 3-8 cycles)",
               "train_steps": train_steps, "train_ms": round(float(nst.train_ms), 1),
               "psnr_untrained_min": round(float(psnr0.min()), 2),
               "psnr_untrained_median": round(float(np.median(psnr0)), 2),
               "psnr_min": round(float(psnr.min()), 2), "psnr_median": round(float(np.median(psnr)), 2),
               "psnr_gain_median_db": round(float(np.median(psnr - psnr0)), 2),
               "materials_decoded": int(nst.n_active), "materials_fallback": int(nst.n_materials - nst.n_active),
               "neural_ms_per_frame": round(ms["neural"], 5), "plain_ms_per_frame": round(ms["plain"], 5),
               "shade_ms": {"plain": round(float(st_plain[3]), 5), "neural": round(float(st_neural[3]), 5)},
               "frame_psnr_neural_vs_plain_db": round(float(10 * np.log10(1.0 / np.mean(diff ** 2))), 2)}
    finally:
        r.close()
    out["ablation_low_res_maps"] = low_res_ablation(scene, col_plain, covered, args)
    scene.close()
    return out


def rsy_decode_timing(scene, seconds=2.0):
    """The workload written as a .rsy file (rosy_amd.asset.from_scene) and decoded by rsy_asset_decode
    from host memory, repeated for a bounded time: MB of file and host GB/s (one thread)."""
    from rosy_amd import asset
    data = asset.encode(asset.from_scene(scene))
    n = 0
    t0 = time.perf_counter()
    while True:
        asset.open_bytes(data).close()
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    ms = (time.perf_counter() - t0) * 1e3 / n
    return {"file_mb": round(len(data) / 1e6, 2), "ms": round(ms, 3), "host_gbs": round(len(data) / (ms * 1e-3) / 1e9, 2),
            "decodes": n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--texture-only", action="store_true",
                    help="only the texture pipeline: BC7 upload with full chains and the plain frame of the decoder "
                         "workload with every map's mip chain (trilinear sampling); prints its JSON")
    ap.add_argument("--frame-only", action="store_true",
                    help="only the decoder workload's frames (no output stage, 4-sample frames, texture pipeline, "
                         "ablation or other workloads): the run the committed profiles are taken of")
    ap.add_argument("--workload", type=int, default=3, choices=[0, 1, 2, 3, 4, 5],
                    help="3 (default): the decoder workload the metric is quoted on; 0: one 2,048-triangle sphere "
                         "at 640x360; 1: scene-scale plain; "
                         "2: 1M triangles, 2560x1440 + 4 cascades; 4: 2M triangles, 4 views + 4 cascades + decoder; "
                         "5: the decoder scene with 16 translucent materials (blended pass), plain shading")
    ap.add_argument("--check", action="store_true",
                    help="workloads 0, 1, 2, 4, 5: compare every view's visibility buffer / shadow map with the "
                         "oracle and every camera view's colour with the FP64 oracle (5: the blended frame's)")
    ap.add_argument("--tile", type=int, default=0, choices=[0, 8, 16, 32, 64],
                    help="bin size in pixels; 0 (default): the autotuner picks it")
    ap.add_argument("--msaa", type=int, default=1, choices=[1, 4],
                    help="workloads 0, 1, 2, 4, 5: samples per camera pixel (4: the reference's 4-sample main pass, "
                         "bin size 32)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 times): the reference kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this frame path is built and measured on a single MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mb:
        import torch
        n = args.copy_probe_mb << 20
        a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "copies": 5}), flush=True)
        return

    if args.texture_only:
        print(json.dumps(texture_timing(args)), flush=True)
        return

    if args.workload != 3:
        run_views(args, rooflines=args.check)
        return

    import numpy as np

    import rosy_amd
    from rosy_amd import abi

    t_gen = time.time()
    scene = rosy_amd.Scene(3, seed=args.seed)
    log(f"scene generated in {time.time() - t_gen:.1f}s: {scene.n_triangles} triangles, "
        f"{scene.s.n_materials} materials")
    r = rosy_amd.Renderer(device=0)
    r.load(scene)
    plain = scene.frame_desc(abi.SHADE_PLAIN)
    neural = scene.frame_desc(abi.SHADE_NEURAL)
    tile = choose_tile(r, scene, plain, args.tile)

    # decoder: train on the GPU, measure PSNR, activate the materials at >= 40 dB
    r.neural_init(decay_steps=args.train_steps)
    psnr0 = r.neural_evaluate()
    loss = r.neural_train(args.train_steps)
    psnr = r.neural_evaluate()
    nst = r.neural_stats()
    log(f"decoder: {args.train_steps} steps in {nst.train_ms:.0f} ms, loss {loss[0]:.2e} -> {loss[-1]:.2e}, "
        f"PSNR min/median {psnr.min():.2f}/{np.median(psnr):.2f} dB (untrained {psnr0.min():.2f}/"
        f"{np.median(psnr0):.2f}), {nst.n_active}/{nst.n_materials} decoded")

    # per-kernel durations of both forms (eager frames, HIP events between the stages)
    st_plain = stage_medians(r, plain)
    st_neural = stage_medians(r, neural)
    stp = r.stats()
    vis = r.visibility(0)
    shadow = r.shadow_map(0)
    col_neural = r.color(0)
    sf_neural = r.shadow_factor(0)
    r.render(plain)
    col_plain = r.color(0)
    sf_plain = r.shadow_factor(0)
    r.set_debug_flags(0)

    # timed region: warmup replays, then exactly `steps` replays of each captured frame graph
    r.render_graph(plain, args.warmup)
    r.render_graph(plain, args.steps)
    plain_ms = float(r.stats().gpu_ms)
    r.render_graph(neural, args.warmup)
    t0 = time.perf_counter()
    r.render_graph(neural, args.steps)  # synchronises at the end
    wall = (time.perf_counter() - t0) * 1e3 / args.steps
    ms = float(r.stats().gpu_ms)
    log(f"graph frame: neural {ms:.4f} ms, plain {plain_ms:.4f} ms (device); neural wall {wall:.4f} ms")

    output_stage = rsy_decode = msaa = texture_pipeline = ablation = fine = None
    if not args.frame_only:
        # output stage of the neural frame (draw image + sRGB blit) and the .rsy reader on this workload
        output_stage = [present_timing(r, 1920, 1080), present_timing(r, 2560, 1440)]
        rsy_decode = rsy_decode_timing(scene)
        log(f"output stage {output_stage}; .rsy decode {rsy_decode}")
        # the same frames with the reference's 4-sample main pass
        msaa = msaa_timing(r, scene, plain, neural, args, check=not args.no_cpu)
        log(f"4-sample frames {msaa}")
        # the texture pipeline: BC7 upload with full chains, and the plain frame with mip chains
        texture_pipeline = texture_timing(args)
        log(f"texture pipeline {texture_pipeline}")
        # what the network earns: the low-resolution-map ablation on this workload, and the same
        # study on content with detail above the latent grid's Nyquist rate
        ablation = low_res_ablation(scene, col_plain, vis != 0, args)
        log(f"low-resolution maps ablation {ablation}")
        fine = decoder_study(args, 1, args.train_steps)
        log(f"fine-detail decoder workload {fine}")
    # the batched-view and combined workloads at their full sizes (every view checked against the
    # oracle, per-stage rooflines)
    larger = {}
    if not args.no_cpu and not args.frame_only:
        import argparse as _ap
        sub = _ap.Namespace(**vars(args))
        sub.check, sub.msaa, sub.tile = True, 1, 0
        for wl in (2, 4):
            o = run_views(sub, workload=wl, emit=False, rooflines=True)
            larger[f"workload_{wl}"] = {k: o[k] for k in ("value", "plain_ms_per_frame", "mtris_per_s", "stage_ms",
                                                           "tile_pairs", "decoder", "bit_exact_vs_oracle",
                                                           "stage_rooflines", "pixels_per_frame")}
            larger[f"workload_{wl}"]["config"] = o["config"]["workload"]
            larger[f"workload_{wl}"]["bin_size"] = o["config"]["bin_size"]

    cam = scene.camera_views()[0]
    cas = scene.cascades()[0]
    raster_ms = float(st_neural[2])
    copy_gbs = copy_ceiling_gbs()

    # decoder MFMA work: decoded pixels x 1088 FLOPs over the neural shading kernel
    active = np.zeros(scene.s.n_materials + 1, bool)
    r_active = psnr >= 40.0
    active[:scene.s.n_materials] = r_active
    surf = [d for d in scene.surfaces() if not d.blended]
    tri_mat = np.concatenate([np.full(d.index_count // 3, min(int(d.material_index), scene.s.n_materials))
                              for d in surf])
    covered = vis != 0
    tri = (vis[covered] & 0xffffffff).astype(np.int64) - 1
    decoded_px = int(active[tri_mat[tri]].sum())
    dec_tflops = decoded_px * DECODER_FLOPS_PER_PIXEL / (float(st_neural[3]) * 1e-3) / 1e12
    decoder = {"decoded_pixels": decoded_px, "flops_per_pixel": DECODER_FLOPS_PER_PIXEL,
               "achieved_tflops": round(dec_tflops, 3), "peak_tflops": MFMA_FP16_PEAK_TFLOPS,
               "mfma_frac": round(dec_tflops / MFMA_FP16_PEAK_TFLOPS, 6),
               "train_ms": round(float(nst.train_ms), 1), "train_steps": args.train_steps,
               "psnr_min": round(float(psnr.min()), 2), "psnr_median": round(float(np.median(psnr)), 2),
               "psnr_untrained_median": round(float(np.median(psnr0)), 2),
               "materials_decoded": int(nst.n_active), "materials": int(nst.n_materials)}

    cpu = None
    err_plain = err_neural = None
    bit_exact = None
    roofline = kernels = pcf_ties = None
    if not args.no_cpu:
        import oracle
        threads = oracle.threads()
        rs = scene.s.raster_state
        n = 0
        t0 = time.perf_counter()
        while True:
            sm, _ = oracle.raster_view(scene, cas.light_view_projection[:], cas.size, cas.size, True, rs, threads)
            ov, _ = oracle.raster_view(scene, cam.view_projection[:], cam.width, cam.height, False, rs, threads)
            oracle.shade_view(scene, cam, [cas], [sm], ov, n_threads=threads)
            n += 1
            if time.perf_counter() - t0 >= args.cpu_seconds or n >= 50:
                break
        cpu_ms = (time.perf_counter() - t0) * 1e3 / n
        cpu = {"value": round(cpu_ms, 2), "unit": "ms/frame", "cores": threads, "kind": "port",
               "sample": f"{n} whole plain frames of the same workload (shadow raster + camera raster + FP64 "
                         f"shading)"}
        bit_exact = bool(np.array_equal(ov, vis) and np.array_equal(sm.view(np.uint32), shadow.view(np.uint32)))
        lat, _ = r.neural_params()
        params = oracle.NeuralParams(np.arange(scene.s.n_materials), lat.astype(np.float16).astype(np.float32),
                                     r.neural_inference_mlp(), r_active.astype(np.uint32))
        err_plain, ties_plain, away_plain = colour_error(col_plain, sf_plain, lambda ov_sf: oracle.shade_view(
            scene, cam, [cas], [sm], ov, override_shadow=ov_sf, n_threads=threads))
        err_neural, ties_neural, away_neural = colour_error(col_neural, sf_neural, lambda ov_sf: oracle.shade_view(
            scene, cam, [cas], [sm], ov, override_shadow=ov_sf, n_threads=threads, neural=params))
        pcf_ties = {"plain": ties_plain, "neural": ties_neural,
                    "differing_away_from_a_tie": away_plain + away_neural}
        # algorithmic bytes of the shading kernels (the oracle's footprint of these frames) and of the
        # geometry, binning and raster kernels
        P = cam.width * cam.height
        fp_plain = oracle.shade_footprint(scene, cam, [cas], [sm], ov, n_threads=threads)
        fp_neural = oracle.shade_footprint(scene, cam, [cas], [sm], ov, neural=params, n_threads=threads)
        b_plain, parts_plain = shade_bytes(fp_plain, P)
        b_neural, parts_neural = shade_bytes(fp_neural, P)
        roofline = hbm_line("k_shade<neural> (deferred Cook-Torrance + PCF shading, decoder for active materials)",
                            b_neural, float(st_neural[3]),
                            {"traffic": measured_traffic("k_shade<neural>", tile), "bytes": parts_neural,
                             "footprint": fp_neural, "measured_copy_gbs": copy_gbs})
        views = [(cas.light_view_projection[:], cas.size, cas.size, True),
                 (cam.view_projection[:], cam.width, cam.height, False)]
        b_raster = raster_bytes(scene, views, rs, tile, int(stp.tile_pairs), int(stp.triangles_setup))
        b_geo = geo_bytes(scene, [(w, h, sh) for _, w, h, sh in views], int(stp.tile_pairs))
        stage_ms_sum = float(st_neural[0] + st_neural[1] + st_neural[2])
        kernels = {
            "k_shade_plain": hbm_line("k_shade<plain>", b_plain, float(st_plain[3]),
                                      {"traffic": measured_traffic("k_shade<plain>", tile), "bytes": parts_plain,
                                       "footprint": fp_plain}),
            "k_raster": hbm_line("k_raster (camera + shadow view in one launch)", b_raster, raster_ms,
                                 {"traffic": measured_traffic("k_raster", tile)}),
            "geometry_binning_raster": hbm_line("transform, setup, clip, bin scan, scatter and raster kernels",
                                                b_geo, stage_ms_sum),
        }
    frame_psnr = float(10 * np.log10(1.0 / np.mean((col_neural[covered][:, :3].astype(np.float64) -
                                                    col_plain[covered][:, :3]) ** 2)))
    out = {
        "metric": METRIC,
        "value": round(ms, 5),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "decoder config: 120 meshes, 262,144 triangles, 1920x1080, 2048^2 shadow map, "
                               "64 materials x 1024^2 RGBA8+RG8+RG8 (noise+stripes+blobs), neural-material shading "
                               "(256^2 x 8 FP16 latents + 10-32-7 MLP, FP16 MFMA), plain-map fallback below 40 dB",
                   "seed": args.seed, "bin_size": tile},
        "plain_ms_per_frame": round(plain_ms, 5),
        "shade_ms": {"plain": round(float(st_plain[3]), 5), "neural": round(float(st_neural[3]), 5)},
        "shading_speedup": round(float(st_plain[3]) / float(st_neural[3]), 3),
        "roofline": roofline,
        "kernels": kernels,
        "decoder": decoder,
        "output_stage": output_stage,
        "msaa4": msaa,
        "texture_pipeline": texture_pipeline,
        "rsy_decode": rsy_decode,
        "cpu_baseline": cpu,
        "wall_ms_per_step": round(wall, 5),
        "mtris_per_s": round(2 * scene.n_triangles / (ms * 1e-3) / 1e6, 1),
        "stage_ms": {"geometry": round(float(st_neural[0]), 5), "binning": round(float(st_neural[1]), 5),
                     "raster": round(raster_ms, 5), "shade": round(float(st_neural[3]), 5)},
        "tile_pairs": int(stp.tile_pairs),
        "max_abs_err_vs_fp64": {"plain": err_plain, "neural": err_neural},
        "pcf_compare_ties": pcf_ties,
        "frame_psnr_neural_vs_plain_db": round(frame_psnr, 2),
        "ablation_low_res_maps": ablation,
        "decoder_fine_detail": fine,
        "larger_workloads": larger,
        "bit_exact_vs_oracle": bit_exact,
    }
    print(json.dumps(out), flush=True)
    r.close()
    scene.close()


WORKLOADS = {
    0: "small: one octahedral geodesic sphere of 2,048 triangles, 640x360, 1024^2 shadow map with 3x3 PCF, one "
       "material of 256^2 maps, plain shading",
    1: "scene-scale: 120 meshes, 262,144 triangles, 1920x1080, 2048^2 shadow map, 16 materials x 2048^2 maps, "
       "plain shading",
    2: "batched views: 1,048,576 triangles (106 meshes), 2560x1440 camera + 4 cascades 2048^2 (splits 5/15/40/100), "
       "16 materials x 2048^2 maps, plain shading, one launch per stage over all 5 views",
    4: "combined: 2,097,152 triangles (116 meshes), cameras 3840x2160 + 1920x1080 + 1280x720 + 640x360 + 4 cascades "
       "2048^2, 64 materials x 1024^2 maps, neural-material shading (FP8 E4M3 decoder weights, FP32 accumulation) "
       "with plain fallback below 40 dB",
    5: "blended: the decoder scene (120 meshes, 262,144 triangles, 1920x1080, 2048^2 shadow map, 64 materials x "
       "1024^2 maps) with materials 0-15 translucent (8 BLEND at alpha 0.7 x a sin-product texture alpha, 8 MASK at "
       "cutoff 0.5) drawn by the blended pass in primitive order, plain shading",
}


def run_views(args, workload=None, emit=True, rooflines=False):
    """Frames of the multi-view workloads: every view of the frame in one launch per stage, replayed
    from a captured graph; value = device ms per frame (neural form where the workload has it).  With
    `rooflines`, the algorithmic bytes of the geometry + binning + raster stage (B_geo) and of the
    shading kernel (the oracle's footprint of every camera view) against their measured times."""
    import numpy as np

    import rosy_amd
    from rosy_amd import abi

    wl = args.workload if workload is None else workload
    t_gen = time.time()
    scene = (rosy_amd.Scene(3, seed=args.seed, n_translucent=16) if wl == 5 else rosy_amd.Scene(wl, seed=args.seed))
    cams, cascs = scene.camera_views(), scene.cascades()
    log(f"workload {wl}: {scene.n_triangles} triangles, {len(cams)} camera views, {len(cascs)} cascades "
        f"(generated in {time.time() - t_gen:.1f}s)")
    r = rosy_amd.Renderer(device=0)
    r.load(scene)
    plain = scene.frame_desc(abi.SHADE_PLAIN)
    if args.msaa == 4:
        import ctypes as C
        rs = abi.RasterState()
        C.memmove(C.byref(rs), C.byref(scene.s.raster_state), C.sizeof(rs))
        rs.msaa_samples, rs.tile_size = 4, 32
        r.update(scene.s.scene_data, rs)
        tile = 32
    else:
        tile = choose_tile(r, scene, plain, args.tile)
    rs_used = scene.s.raster_state if args.msaa == 1 else rs
    frame = plain
    dec = None
    if wl == 4:
        r.neural_init(decay_steps=args.train_steps, weights=abi.NEURAL_WEIGHTS_FP8)
        r.neural_train(args.train_steps)
        psnr = r.neural_evaluate()
        nst = r.neural_stats()
        dec = {"psnr_min": round(float(psnr.min()), 2), "psnr_median": round(float(np.median(psnr)), 2),
               "materials_decoded": int(nst.n_active), "materials": int(nst.n_materials),
               "train_ms": round(float(nst.train_ms), 1), "weights": "fp8 e4m3 (per-layer scale)"}
        frame = scene.frame_desc(abi.SHADE_NEURAL)
    st = stage_medians(r, frame)
    fst = r.stats()
    exact = None
    blend_check = None
    colour_err = None
    if args.check:
        import oracle
        rs = rs_used
        exact = True
        for ci, cs in enumerate(cascs):
            o, _ = oracle.raster_view(scene, cs.light_view_projection[:], cs.size, cs.size, True, rs)
            exact &= bool(np.array_equal(r.shadow_map(ci).view(np.uint32), o.view(np.uint32)))
        for vi, cv in enumerate(cams):
            if args.msaa == 4:
                o = oracle.raster_view_msaa(scene, cv.view_projection[:], cv.width, cv.height, rs)
            else:
                o, _ = oracle.raster_view(scene, cv.view_projection[:], cv.width, cv.height, False, rs)
            exact &= bool(np.array_equal(r.visibility(vi), o))
        log(f"visibility + shadow maps bit-exact vs oracle: {exact}")
        if wl != 5 and args.msaa == 1:
            # the measured frame's colour in every camera view against the FP64 oracle (PCF tie rule)
            maps = [r.shadow_map(ci) for ci in range(len(cascs))]
            params = None
            if wl == 4:
                lat, _ = r.neural_params()
                params = oracle.NeuralParams(np.arange(scene.s.n_materials), lat.astype(np.float16).astype(np.float32),
                                             r.neural_inference_mlp(), (psnr >= 40.0).astype(np.uint32))
            colour_err = []
            for vi, cv in enumerate(cams):
                e, ties, away = colour_error(r.color(vi), r.shadow_factor(vi), lambda ov_sf: oracle.shade_view(
                    scene, cv, cascs, maps, r.visibility(vi), override_shadow=ov_sf, neural=params))
                colour_err.append({"view": f"{cv.width}x{cv.height}", "max_abs_err_vs_fp64": e, "pcf_ties": ties,
                                   "pcf_differing_away_from_a_tie": away})
            log(f"colour vs FP64 oracle: {colour_err}")
        if wl == 5:
            vis = r.visibility(0)
            maps = [r.shadow_map(0)]
            if args.msaa == 4:
                want, _, counts, _ = oracle.msaa_view(scene, cams[0], cascs, maps, vis, raster_state=rs, blended=True,
                                                      max_layers=64)
                count = counts.max(axis=0)
            else:
                base, _, _ = oracle.shade_view(scene, cams[0], cascs, maps, vis)
                want, count = oracle.blend_view(scene, cams[0], cascs, maps, vis, base, max_layers=64)
            blend_check = {"max_abs_err_vs_fp64": float(np.abs(r.color(0).astype(np.float64) - want).max()),
                           "pixels_with_layers": int((count > 0).sum()), "max_layers": int(count.max())}
            log(f"blended frame vs FP64 oracle: {blend_check}")
    stage_rooflines = None
    if rooflines:
        import oracle
        rs = rs_used
        maps = [r.shadow_map(ci) for ci in range(len(cascs))]
        params = None
        if wl == 4:
            lat, _ = r.neural_params()
            active = (psnr >= 40.0).astype(np.uint32)
            params = oracle.NeuralParams(np.arange(scene.s.n_materials), lat.astype(np.float16).astype(np.float32),
                                         r.neural_inference_mlp(), active)
        b_shade, parts = 0, {}
        for vi, cv in enumerate(cams):
            fp = oracle.shade_footprint(scene, cv, cascs, maps, r.visibility(vi), neural=params)
            b, pr = shade_bytes(fp, cv.width * cv.height)
            b_shade += b
            for k, v in pr.items():
                parts[k] = parts.get(k, 0) + v
        views_px = [(c.width, c.height, False) for c in cams] + [(c.size, c.size, True) for c in cascs]
        b_geo = geo_bytes(scene, views_px, int(fst.tile_pairs))
        stage_rooflines = {
            "geometry_binning_raster": hbm_line("transform, setup, clip, bin scan, scatter and raster kernels", b_geo,
                                                float(st[0] + st[1] + st[2])),
            "shade": hbm_line("k_shade<neural>" if wl == 4 else "k_shade<plain>", b_shade, float(st[3]),
                              {"bytes": parts}),
        }
        log(f"stage rooflines {stage_rooflines}")
    r.set_debug_flags(0)
    r.render_graph(plain, args.warmup)
    r.render_graph(plain, args.steps)
    plain_ms = float(r.stats().gpu_ms)
    r.render_graph(frame, args.warmup)
    r.render_graph(frame, args.steps)
    ms = float(r.stats().gpu_ms)
    views = len(cams) + len(cascs)
    out = {"metric": "ms/frame (all views of the frame)", "value": round(ms, 5), "unit": "ms/frame", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": WORKLOADS[wl], "seed": args.seed, "bin_size": tile,
                      "samples_per_pixel": args.msaa},
           "plain_ms_per_frame": round(plain_ms, 5),
           "mtris_per_s": round(views * scene.n_triangles / (ms * 1e-3) / 1e6, 1),
           "pixels_per_frame": int(sum(c.width * c.height for c in cams) + sum(c.size * c.size for c in cascs)),
           "stage_ms": {"geometry": round(float(st[0]), 5), "binning": round(float(st[1]), 5),
                        "raster": round(float(st[2]), 5), "shade": round(float(st[3]), 5),
                        "blended_pass": round(float(st[4]), 5)},
           "tile_pairs": int(fst.tile_pairs), "triangles_setup": int(fst.triangles_setup),
           "triangles_clipped": int(fst.triangles_clipped), "decoder": dec, "bit_exact_vs_oracle": exact,
           "colour": colour_err, "blended": blend_check, "stage_rooflines": stage_rooflines}
    if emit:
        print(json.dumps(out), flush=True)
    r.close()
    scene.close()
    return out


if __name__ == "__main__":
    main()
