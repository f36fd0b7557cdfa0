"""GPU parity: the gfx950 kernels against the CPU oracle (oracle/).

Bar: LBVH stages (Morton codes, stable sort, Karras topology, refit bounds)
and closest-hit (t, primitive id, step and leaf-test counts) bit-exact;
images per pixel within relMSE < 1e-4 (mean of (a-b)^2/(b^2+1e-3)), and the
three GPU integrators bit-identical to each other (the reference's own
invariant, README.md:17, :136, :154).
"""
import numpy as np
import pytest

from conftest import rel_mse

pytestmark = pytest.mark.gpu

LBVH_KEYS = ("morton", "sorted_keys", "sorted_ids", "left", "right", "left_is_leaf", "right_is_leaf", "parent",
             "leaf_parent")


def _bounds_equal(a, b):
    return np.array_equal(a["min"], b["min"]) and np.array_equal(a["max"], b["max"])


def _check_lbvh(mcpt, ctx, oracle, hot, cold, mats):
    print(f"lbvh parity n={len(hot)}", flush=True)
    ref = oracle.Scene(hot, cold, mats).dump()
    scene = ctx.scene(hot, cold, mats)
    got = scene.build(dump=True)
    for k in LBVH_KEYS:
        assert np.array_equal(got[k], ref[k]), k
    assert _bounds_equal(got["internal_bounds"], ref["internal_bounds"])
    assert _bounds_equal(got["leaf_bounds"], ref["leaf_bounds"])
    assert _bounds_equal(got["scene_bounds"], ref["scene_bounds"])
    assert got["max_depth"] == ref["max_depth"]
    return scene


def test_lbvh_cornell(mcpt, ctx, oracle):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    _check_lbvh(mcpt, ctx, oracle, hot, cold, mats).close()


@pytest.mark.parametrize("n", [1, 2, 3, 17, 4097, 65536])
def test_lbvh_soup_sizes(mcpt, ctx, oracle, n):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=n, param=n)
    if n <= 16:  # drop the emitters to get exactly n triangles
        hot, cold = hot[:n], cold[:n]
    _check_lbvh(mcpt, ctx, oracle, hot, cold, mats).close()


def test_lbvh_duplicate_keys(mcpt, ctx, oracle):
    # all centroids equal -> every Morton code equal: the tie rule (index
    # bits) alone shapes the tree
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=3, param=1000)
    c = (hot["v0"] + hot["v1"] + hot["v2"]) / 3
    for k in ("v0", "v1", "v2"):
        hot[k] = hot[k] - c + np.float32(0.5)
    _check_lbvh(mcpt, ctx, oracle, hot, cold, mats).close()


def test_lbvh_full_soup(mcpt, ctx, oracle):
    """the 1,048,592-triangle incoherent mesh at full size"""
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=1)
    assert len(hot) == 1048592
    _check_lbvh(mcpt, ctx, oracle, hot, cold, mats).close()


def _reference_leaf_order(d):
    """Each leaf's position in the reference's leaf-test order, from a
    left-first pre-order walk of the oracle's tree in which a node's leaf
    children are tested when it is visited, left then right
    (BVH::traverseRay, src/path_tracer.cu:1102-1147)."""
    L, R, Ll, Rl = d["left"], d["right"], d["left_is_leaf"], d["right_is_leaf"]
    rank = np.zeros(len(d["leaf_parent"]), np.int64)
    pos, stack = 0, [0] if len(L) else []
    while stack:
        p = stack.pop()
        for c, is_leaf in ((L[p], Ll[p]), (R[p], Rl[p])):
            if is_leaf:
                rank[c] = pos
                pos += 1
        if not Rl[p]:
            stack.append(R[p])
        if not Ll[p]:
            stack.append(L[p])
    return rank


def _box(b):
    return np.concatenate([b["min"], b["max"]]).astype(np.float32)


def _check_wide(mcpt, scene, ref):
    """The four-wide tree against the oracle's binary tree: wide node i is
    the i-th even-level binary node in (level, index) order, so every wide
    node's inner children have consecutive indices in slot order; its
    children are the binary node's leaf
    children and its inner children's children in left-to-right order, every
    decoded box contains the exact one and exceeds it by less than one
    quantization step per side, the scale per axis is the smallest power of
    two (>= 2^-126) that spans the node in 255 steps, and the leaf ranks
    follow the reference's leaf-test order."""
    nodes, tie = scene.dump_wide()
    L, R, Ll, Rl = ref["left"], ref["right"], ref["left_is_leaf"], ref["right_is_leaf"]
    ni, n = len(L), len(ref["leaf_parent"])
    order = np.argsort(_reference_leaf_order(ref), kind="stable")
    assert len(np.unique(tie)) == n and np.all(np.diff(tie[order]) > 0)
    if ni == 0:
        assert len(nodes) == 1 and list(nodes[0]["child"][:2]) == [-1, -1]
        return
    level = np.full(ni, -1)
    level[0] = 0
    stack = [0]
    while stack:
        p = stack.pop()
        for c, is_leaf in ((L[p], Ll[p]), (R[p], Rl[p])):
            if not is_leaf:
                level[c] = level[p] + 1
                stack.append(c)
    even = np.flatnonzero(level % 2 == 0)
    even = even[np.lexsort((even, level[even]))]
    assert len(nodes) == len(even)
    widx = np.full(ni, -1)
    widx[even] = np.arange(len(even))
    exp_ref = np.zeros((len(even), 4), np.int64)
    exp_box = np.zeros((len(even), 4, 6), np.float32)
    cnt = np.zeros(len(even), np.int64)
    for i, p in enumerate(even):
        for c, is_leaf in ((L[p], Ll[p]), (R[p], Rl[p])):
            kids = [(c, True)] if is_leaf else [(L[c], Ll[c]), (R[c], Rl[c])]
            for g, g_leaf in kids:
                exp_ref[i, cnt[i]] = ~g if g_leaf else widx[g]
                exp_box[i, cnt[i]] = _box(ref["leaf_bounds"][g] if g_leaf else ref["internal_bounds"][g])
                cnt[i] += 1
    assert np.array_equal(nodes["meta"] >> 24, cnt)
    valid = np.arange(4)[None, :] < cnt[:, None]
    assert np.array_equal(np.where(valid, nodes["child"], 0), np.where(valid, exp_ref, 0))
    # inner children consecutive, in slot order (the walk's pops rely on it)
    for i in range(len(even)):
        inner = [int(r) for r in nodes["child"][i, :cnt[i]] if r >= 0]
        assert inner == list(range(inner[0], inner[0] + len(inner))) if inner else True, i
    # every leaf once, every wide node but the root once
    refs = nodes["child"][valid]
    assert np.array_equal(np.sort(~refs[refs < 0]), np.arange(n))
    assert np.array_equal(np.sort(refs[refs >= 0]), np.arange(1, len(even)))
    # decode in float32 exactly as the traversal bounds it
    e = np.stack([(nodes["meta"] >> (8 * a)) & 255 for a in range(3)], 1).astype(np.int64)
    scale = np.ldexp(np.float32(1.0), (e - 127).astype(np.int32)).astype(np.float32)  # (W, 3)
    q = nodes["q"].astype(np.float32)  # (W, 6, 4): lo xyz, hi xyz per child
    org = nodes["origin"]
    lo = org[:, :, None] + q[:, 0:3, :] * scale[:, :, None]
    hi = org[:, :, None] + q[:, 3:6, :] * scale[:, :, None]
    lo, hi = np.transpose(lo, (0, 2, 1)), np.transpose(hi, (0, 2, 1))  # (W, 4, 3)
    emin, emax = exp_box[:, :, 0:3], exp_box[:, :, 3:6]
    v3 = valid[:, :, None]
    assert np.all(~v3 | (lo <= emin)) and np.all(~v3 | (hi >= emax))
    # at most one quantization step outside, plus the half ulp by which the
    # float decode origin + q * 2^e can fall short of an exact multiple (the
    # encoder then steps once more)
    s3 = scale[:, None, :]
    assert np.all(~v3 | (emin - lo <= s3 + np.spacing(emin))) and np.all(~v3 | (hi - emax <= s3 + np.spacing(emax)))
    assert np.array_equal(org, np.where(v3, emin, np.inf).min(axis=1))
    ext = np.where(v3, emax.astype(np.float64), -np.inf).max(axis=1) - org.astype(np.float64)
    assert np.all(255.0 * scale >= ext)
    assert np.all((e == 1) | (255.0 * scale / 2 <= ext))


@pytest.mark.parametrize("n", [1, 2, 3, 17, 4097, 65536])
def test_wide_tree(mcpt, ctx, oracle, n):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=n, param=n)
    if n <= 16:
        hot, cold = hot[:n], cold[:n]
    print(f"wide tree n={len(hot)}", flush=True)
    ref = oracle.Scene(hot, cold, mats).dump()
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    _check_wide(mcpt, scene, ref)
    scene.close()


def test_wide_tree_cornell_and_duplicates(mcpt, ctx, oracle):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    scenes = [(hot, cold, mats)]
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=3, param=1000)
    c = (hot["v0"] + hot["v1"] + hot["v2"]) / 3
    for k in ("v0", "v1", "v2"):
        hot[k] = hot[k] - c + np.float32(0.5)
    scenes.append((hot, cold, mats))
    for hot, cold, mats in scenes:
        ref = oracle.Scene(hot, cold, mats).dump()
        scene = ctx.scene(hot, cold, mats)
        scene.build()
        _check_wide(mcpt, scene, ref)
        scene.close()


ROOM_MTL = """newmtl white
Kd 0.73 0.73 0.73
newmtl red
Kd 0.65 0.05 0.05
newmtl lamp
Ke 15 15 15
"""

ROOM_OBJ = """mtllib room.mtl
v -1 0 -1
v 1 0 -1
v 1 0 1
v -1 0 1
v -1 2 -1
v 1 2 -1
v 1 2 1
v -1 2 1
v -0.3 1.98 -0.3
v 0.3 1.98 -0.3
v 0.3 1.98 0.3
v -0.3 1.98 0.3
v -0.5 0 -0.2
v 0.1 0 -0.4
v 0.3 0 0.2
v -0.3 0 0.4
v -0.5 0.8 -0.2
v 0.1 0.8 -0.4
v 0.3 0.8 0.2
v -0.3 0.8 0.4
usemtl white
f 1 2 3 4
f 5 8 7 6
f 1 5 6 2
usemtl red
f 1 4 8 5
usemtl white
f 2 6 7 3
usemtl lamp
f 9 10 11 12
usemtl white
f 13 14 15 16
f 17 20 19 18
f 13 17 18 14
f 14 18 19 15
f 15 19 20 16
f 16 20 17 13
"""


def test_render_obj_scene(mcpt, ctx, oracle, tmp_path):
    """An OBJ + MTL room loaded by mcpt_load_obj renders like the oracle's
    render of the same records, in every mode."""
    (tmp_path / "room.mtl").write_text(ROOM_MTL)
    (tmp_path / "room.obj").write_text(ROOM_OBJ)
    hot, cold, mats, st = mcpt.load_obj(str(tmp_path / "room.obj"), translation=(0, 0, 0), scale=2.0)
    assert st["n_tris"] == 24 and st["n_materials"] == 4
    W = H = 64
    spp, depth, seed = 4, 8, 3
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    desc.look_from[:] = (0.0, 1.0, 3.5)
    desc.look_at[:] = (0.0, 1.0, 0.0)
    cam = mcpt.camera_init(desc)
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    ref = oracle.Scene(hot, cold, mats)
    ref_img, _, _ = ref.render(oracle.camera_init(desc), spp, depth, seed=seed)
    assert ref_img.mean() > 0.01  # the lamp lights the room
    out = _render_modes(mcpt, ctx, scene, cam, spp, depth, seed)
    base = out[mcpt.MODE_MEGAKERNEL][0]
    for mode, (img, _) in out.items():
        assert np.array_equal(img, base), mode
        assert rel_mse(img, ref_img) < 1e-4, mode
    scene.close()


def test_render_textures(mcpt, ctx, oracle, tmp_path):
    """A room with a checker floor and an image on the back wall renders the
    same on the GPU (every mode) and in the oracle (sampleTexture,
    src/path_tracer.cu:216-234, the same FP32 filter on both sides)."""
    (tmp_path / "room.mtl").write_text(ROOM_MTL)
    text = ROOM_OBJ.replace("usemtl white\nf 1 2 3 4\nf 5 8 7 6\nf 1 5 6 2\n",
                            "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                            "usemtl white\nf 1/1 2/2 3/3 4/4\nf 5 8 7 6\nf 1/1 5/4 6/3 2/2\n")
    (tmp_path / "room.obj").write_text(text)
    hot, cold, mats, st = mcpt.load_obj(str(tmp_path / "room.obj"), scale=2.0)
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    # materials: the floor and back wall (white, used by faces 0-1 and 4-5)
    # get their own copies: floor -> checker, back wall -> image
    mats = np.concatenate([mats, mats[1:2], mats[1:2]])
    mats[-2]["texture"], mats[-1]["texture"] = 1, 2
    cold["material"][0:2] = len(mats) - 2
    cold["material"][4:6] = len(mats) - 1
    W = H = 64
    spp, depth, seed = 4, 6, 9
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    desc.look_from[:] = (0.0, 1.0, 3.5)
    desc.look_at[:] = (0.0, 1.0, 0.0)
    cam = mcpt.camera_init(desc)
    scene = ctx.scene(hot, cold, mats)
    ref = oracle.Scene(hot, cold, mats)
    for target in (scene, ref):
        assert target.add_texture(mcpt.TEX_CHECKER, color_a=(0.9, 0.9, 0.9), color_b=(0.1, 0.3, 0.8), scale=6.0) == 0
        assert target.add_texture(mcpt.TEX_IMAGE, image=img) == 1
    scene.build()
    ref_img, _, _ = ref.render(oracle.camera_init(desc), spp, depth, seed=seed)
    out = _render_modes(mcpt, ctx, scene, cam, spp, depth, seed)
    base = out[mcpt.MODE_MEGAKERNEL][0]
    for mode, (im, _) in out.items():
        assert np.array_equal(im, base), mode
        assert rel_mse(im, ref_img) < 1e-4, mode
        print(f"textures mode {mode}: identical to the oracle {np.mean(im == ref_img):.5f}", flush=True)
    # a material naming a texture the scene lacks is refused
    lone = ctx.scene(hot, cold, mats)
    lone.build()
    film = ctx.film(W, H)
    with pytest.raises(mcpt.McptError):
        lone.render(cam, film, spp=1, max_depth=2)
    film.close()
    lone.close()
    scene.close()


def test_render_obj_map_kd(mcpt, ctx, oracle, tmp_path):
    """A map_Kd image read by the OBJ loader (load_textures) and handed to
    the scene renders the same on the GPU and in the oracle."""
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (16, 24, 3), dtype=np.uint8)
    (tmp_path / "poster.ppm").write_bytes(b"P6\n24 16\n255\n" + img.tobytes())
    (tmp_path / "room.mtl").write_text(ROOM_MTL + "newmtl poster\nKd 0.5 0.5 0.5\nmap_Kd poster.png\n")
    text = ROOM_OBJ.replace("usemtl white\nf 1 2 3 4\nf 5 8 7 6\nf 1 5 6 2\n",
                            "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                            "usemtl white\nf 1 2 3 4\nf 5 8 7 6\nusemtl poster\nf 1/1 5/4 6/3 2/2\n")
    (tmp_path / "room.obj").write_text(text)
    hot, cold, mats, st, images = mcpt.load_obj(str(tmp_path / "room.obj"), scale=2.0, load_textures=True)
    assert st["textures_loaded"] == 1 and len(images) == 1 and np.count_nonzero(mats["texture"]) == 1
    W = H = 48
    spp, depth, seed = 4, 5, 21
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    desc.look_from[:] = (0.0, 1.0, 3.5)
    desc.look_at[:] = (0.0, 1.0, 0.0)
    cam = mcpt.camera_init(desc)
    scene = ctx.scene(hot, cold, mats)
    ref = oracle.Scene(hot, cold, mats)
    for target in (scene, ref):
        assert target.add_texture(mcpt.TEX_IMAGE, image=images[0]) == 0
    scene.build()
    ref_img, _, _ = ref.render(oracle.camera_init(desc), spp, depth, seed=seed)
    out = _render_modes(mcpt, ctx, scene, cam, spp, depth, seed)
    for mode, (im, _) in out.items():
        assert rel_mse(im, ref_img) < 1e-4, mode
        print(f"map_Kd mode {mode}: identical to the oracle {np.mean(im == ref_img):.5f}", flush=True)
    scene.close()


def _rays(n, seed, origin_box=(0.0, 1.0), tmax=1e30):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, mcpt_ray_dtype())
    r["origin"] = rng.uniform(origin_box[0], origin_box[1], (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["dir"] = d.astype(np.float32)
    r["tmax"] = np.float32(tmax)
    return r


def mcpt_ray_dtype():
    return np.dtype([("origin", "<f4", 3), ("time", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])


@pytest.mark.parametrize("n_tris", [4096, 65536])
def test_trace_closest_and_any(mcpt, ctx, oracle, n_tris):
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=11, param=n_tris)
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    ref = oracle.Scene(hot, cold, mats)
    rays = _rays(1 << 16, n_tris, (-0.2, 1.2))
    print(f"trace parity n_tris={n_tris}", flush=True)
    # with counts: the reference's visit order, counts must match too
    hits, steps, leaves = scene.trace_closest(rays, counts=True)
    rh, rs, rl = ref.trace_closest(rays)
    assert np.array_equal(hits["prim_index"], rh["prim_index"])
    assert np.array_equal(hits["t"].view(np.uint32), rh["t"].view(np.uint32))
    assert np.array_equal(hits["did_hit"], rh["did_hit"])
    assert np.array_equal(steps, rs) and np.array_equal(leaves, rl)
    # without: near-first order, same hits bit for bit
    fast = scene.trace_closest(rays)
    assert np.array_equal(fast, hits)
    assert hits["did_hit"].mean() > (0.05 if n_tris >= 65536 else 0.005)
    # any-hit with a finite limit
    rays2 = rays.copy()
    rays2["tmax"] = np.float32(0.05)
    occ = scene.trace_any(rays2)
    assert np.array_equal(occ, ref.trace_any(rays2))
    assert 0 < occ.mean() < 1
    scene.close()


def test_trace_exact_ties(mcpt, ctx, oracle):
    """A flat grid of 2-triangle quads hit exactly on shared edges and
    vertices: many triangles share the same t, so the winner is decided by
    the reference's test order alone (near-first traversal must reproduce
    it)."""
    g = 24
    tris = []
    for i in range(g):
        for j in range(g):
            x0, x1, y0, y1 = i / g, (i + 1) / g, j / g, (j + 1) / g
            if (i + j) % 2:
                tris += [((x0, y0, 0), (x1, y0, 0), (x1, y1, 0)), ((x0, y0, 0), (x1, y1, 0), (x0, y1, 0))]
            else:
                tris += [((x0, y0, 0), (x1, y0, 0), (x0, y1, 0)), ((x1, y0, 0), (x1, y1, 0), (x0, y1, 0))]
    hot = np.zeros(len(tris), mcpt.HOT_DTYPE)
    for k, t in enumerate(tris):
        hot[k]["v0"], hot[k]["v1"], hot[k]["v2"] = t
    cold = np.zeros(len(tris), mcpt.COLD_DTYPE)
    mats = np.zeros(1, mcpt.MAT_DTYPE)
    pts = [(i / g, j / g) for i in range(g + 1) for j in range(g + 1)]            # vertices
    pts += [((i + 0.5) / g, j / g) for i in range(g) for j in range(g + 1)]       # edges
    pts += [((i + 0.5) / g, (j + 0.5) / g) for i in range(g) for j in range(g)]   # diagonals
    rays = np.zeros(len(pts), mcpt_ray_dtype())
    rays["origin"] = [(x, y, 1.0) for x, y in pts]
    rays["dir"] = (0.0, 0.0, -1.0)
    rays["tmax"] = 1e30
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    rh, _, _ = oracle.Scene(hot, cold, mats).trace_closest(rays)
    assert rh["did_hit"].sum() > len(pts) // 2
    for counts in (True, False):
        got = scene.trace_closest(rays, counts=counts)
        got = got[0] if counts else got
        assert np.array_equal(got["prim_index"], rh["prim_index"]), counts
        assert np.array_equal(got["t"].view(np.uint32), rh["t"].view(np.uint32))
    scene.close()


def test_trace_axis_parallel(mcpt, ctx, oracle):
    """Directions with zero or denormal components (infinite or huge
    reciprocals, e.g. the centre pixel of an odd-sized image without jitter):
    both visit orders must return the reference's hits."""
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=5, param=20000)
    rng = np.random.default_rng(3)
    n = 6000
    rays = np.zeros(n, mcpt_ray_dtype())
    rays["origin"] = rng.uniform(-0.1, 1.1, (n, 3)).astype(np.float32)
    d = np.zeros((n, 3))
    axis = rng.integers(3, size=n)
    d[np.arange(n), axis] = rng.choice([-1.0, 1.0], n)
    two = np.arange(n) % 3 == 1  # one zero component
    d[two] = rng.normal(size=(two.sum(), 3))
    d[two, axis[two]] = 0.0
    tiny = np.arange(n) % 3 == 2  # denormal components
    d[tiny, (axis[tiny] + 1) % 3] = 1e-39
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["dir"] = d.astype(np.float32)
    rays["tmax"] = 1e30
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    rh, _, _ = oracle.Scene(hot, cold, mats).trace_closest(rays)
    assert rh["did_hit"].sum() > 50
    for counts in (True, False):
        got = scene.trace_closest(rays, counts=counts)
        got = got[0] if counts else got
        assert np.array_equal(got["prim_index"], rh["prim_index"]), counts
        assert np.array_equal(got["t"].view(np.uint32), rh["t"].view(np.uint32))
    rays["tmax"] = 0.3
    assert np.array_equal(scene.trace_any(rays), oracle.Scene(hot, cold, mats).trace_any(rays))
    scene.close()


def _render_modes(mcpt, ctx, scene, cam, spp, depth, seed):
    """Every render mode; the persistent mode both with its persistent shadow
    kernel (the default) and with one lane per shadow ray."""
    out = {}
    runs = [(m, m, 0) for m in (mcpt.MODE_MEGAKERNEL, mcpt.MODE_WAVEFRONT, mcpt.MODE_PERSISTENT)]
    runs.append(("persistent, shadow per ray", mcpt.MODE_PERSISTENT, mcpt.RENDER_SHADOW_PER_RAY))
    for key, mode, flags in runs:
        film = ctx.film(cam.width, cam.height)
        st = scene.render(cam, film, mode=mode, spp=spp, max_depth=depth, seed=seed, flags=flags)
        out[key] = (film.read(), st)
        film.close()
    return out


def test_render_cornell(mcpt, ctx, oracle):
    W = H = 128
    spp, depth, seed = 4, 8, 5
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    cam = mcpt.camera_init(desc)
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    ref_img, _, ref_st = oracle.Scene(hot, cold, mats).render(oracle.camera_init(desc), spp, depth, seed=seed)
    out = _render_modes(mcpt, ctx, scene, cam, spp, depth, seed)
    base = out[mcpt.MODE_MEGAKERNEL][0]
    for mode, (img, st) in out.items():
        print(f"cornell mode {mode}: relMSE {rel_mse(img, ref_img):.3e} identical {np.mean(img == ref_img):.5f}",
              flush=True)
        assert np.array_equal(img, base), f"mode {mode} differs from the megakernel"
        assert rel_mse(img, ref_img) < 1e-4
        assert st["extend_rays"] == ref_st["extend_rays"]
    # samples in flight per pass (4 samples as 3 + 1, and 1 + 1 + 1 + 1): same film
    for batch in (3, 1):
        film = ctx.film(W, H)
        scene.render(cam, film, mode=mcpt.MODE_PERSISTENT, spp=spp, max_depth=depth, seed=seed, batch=batch)
        assert np.array_equal(film.read(), base), batch
        film.close()
    # reference visit order: same image, and the extend kernels' step and
    # leaf-test totals equal the oracle's
    for mode in (mcpt.MODE_WAVEFRONT, mcpt.MODE_PERSISTENT):
        film = ctx.film(W, H)
        st = scene.render(cam, film, mode=mode, spp=spp, max_depth=depth, seed=seed,
                          flags=mcpt.RENDER_COUNT_TRAVERSAL | mcpt.RENDER_REFERENCE_ORDER)
        assert np.array_equal(film.read(), base)
        assert st["extend_steps"] == ref_st["extend_steps"] and st["extend_leaves"] == ref_st["extend_leaves"]
        film.close()
    scene.close()


def test_render_soup_sky(mcpt, ctx, oracle):
    W = H = 96
    spp, depth, seed = 2, 16, 9
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=2, param=65536)
    sky = mcpt.gen_sky(512, 256)
    desc = mcpt.gen_camera(mcpt.GEN_SOUP, W, H)
    cam = mcpt.camera_init(desc)
    scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))
    scene.build()
    ref = oracle.Scene(hot, cold, mats, sky=sky)
    ref_img, _, ref_st = ref.render(oracle.camera_init(desc), spp, depth, seed=seed)
    out = _render_modes(mcpt, ctx, scene, cam, spp, depth, seed)
    base = out[mcpt.MODE_MEGAKERNEL][0]
    for mode, (img, st) in out.items():
        print(f"soup mode {mode}: relMSE {rel_mse(img, ref_img):.3e} identical {np.mean(img == ref_img):.5f}",
              flush=True)
        assert np.array_equal(img, base), f"mode {mode} differs from the megakernel"
        assert rel_mse(img, ref_img) < 1e-4
    scene.close()


@pytest.mark.parametrize("scale,offset", [(1000.0, (4096.0, -8192.0, 12288.0)),
                                          (2000.0, (-20000.0, 9000.0, 25000.0))])
def test_trace_large_coordinates(mcpt, ctx, oracle, scale, offset):
    """The 64k-triangle soup scaled up and moved far from the origin, so that
    coordinates are 1e4-3e4 while the 1e-3 box padding stays the same: both
    visit orders return the reference's hits, for rays inside the mesh and
    primary rays from a camera 3000 units away."""
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=12, param=65536)
    off = np.array(offset)
    for k in ("v0", "v1", "v2"):
        hot[k] = (hot[k].astype(np.float64) * scale + off).astype(np.float32)
    n = 1 << 16
    rng = np.random.default_rng(8)
    rays = np.zeros(n, mcpt_ray_dtype())
    rays["origin"] = (rng.uniform(-0.1, 1.1, (n, 3)) * scale + off).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    far = np.arange(n) % 4 == 0  # from a point 1.5 scene sizes in front, aimed into the mesh
    eye = off + scale * np.array([0.5, 0.5, -1.5])
    tgt = rng.uniform(0.0, 1.0, (far.sum(), 3)) * scale + off
    dd = tgt - eye
    rays["origin"][far] = eye.astype(np.float32)
    rays["dir"][far] = (dd / np.linalg.norm(dd, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.float32(1e30)
    print(f"trace large coordinates scale {scale} offset {offset}", flush=True)
    scene = ctx.scene(hot, cold, mats)
    scene.build()
    ref = oracle.Scene(hot, cold, mats)
    rh, rs, rl = ref.trace_closest(rays)
    assert rh["did_hit"].mean() > 0.1
    hits, steps, leaves = scene.trace_closest(rays, counts=True)
    assert np.array_equal(hits, rh) and np.array_equal(steps, rs) and np.array_equal(leaves, rl)
    fast = scene.trace_closest(rays)
    bad = np.flatnonzero(fast != rh)
    assert bad.size == 0, f"four-wide walk differs on {bad.size} rays, e.g. {bad[:5]}: {fast[bad[:3]]} vs {rh[bad[:3]]}"
    rays["tmax"] = np.float32(0.05 * scale)
    assert np.array_equal(scene.trace_any(rays), ref.trace_any(rays))
    scene.close()
