"""GPU parity of the strand rasterizer against the float64 oracle, in the
two stages the blend-order-sensitive k-buffer calls for:
  1. fragments: the set of (segment, pixel) pairs, their colour before
     packUnorm4x8 (max abs error 1e-4) and depth (1e-6), read through the
     per-node capture; the packed node colour is packUnorm4x8 of it;
  2. resolve: given the identical node lists (atomic insertion order is not
     deterministic), the k-buffer resolve within 2e-3 of float64.
Also: the shadow-map depth pass, PPLL bookkeeping, pool overflow without
out-of-bounds writes, the hybrid LoD (strand coverage x (1 - lod), raymarch
coverage x lod, one shared list), and the full 1080p / 1024^2 / 32M-node
configuration on every segment."""
import numpy as np
import pytest
import torch

from oracle import raster as ORs
from oracle import raymarch as OR
from tests.parity import compare_raymarch
from vkhr_amd import capi, params, synth
from vkhr_amd import device as D

pytestmark = pytest.mark.gpu

FRAG_TOL = 1e-4
RESOLVE_TOL = 2e-3
NULL = 0xFFFFFFFF


def _cm(m):
    return np.asarray(m, np.float64).T.reshape(-1).astype(np.float32)


class Scene:
    def __init__(self, style, res, w, h, device, azimuth=0.0, shadow=1024, segments=True, line_width=1.0):
        self.style, self.res, self.w, self.h = style, res, w, h
        self.xyz = D.to_dev(style.vertices, device)
        off = D.to_dev(style.strand_offsets, device)
        self.idx = D.generate_indices(off)
        self.tan = D.generate_tangents(self.xyz, off)
        self.th = D.to_dev(style.thickness, device)
        self.bounds, self.vol = D.style_volume(style, res, device, segments=segments)
        self.sc = synth.make_scene(self.bounds, w, h, azimuth_deg=azimuth)
        self.cam, self.light = params.camera(self.sc), params.light(self.sc)
        self.sp = params.strand(self.bounds, res, strand_radius=line_width)
        self.rp = params.render(renderer=0)
        self.shadow = shadow
        self.smap = D.shadow_depth(self.xyz, self.idx, _cm(self.sc.light_vp), line_width, (shadow, shadow))

    def np_inputs(self):
        return dict(xyz=self.style.vertices, tan=self.tan.cpu().numpy(), th=self.style.thickness,
                    idx=self.idx.cpu().numpy().view(np.uint32), dens=self.vol["density"].cpu().numpy(),
                    smap=self.smap.cpu().numpy())

    def raster(self, node_count, capture=True):
        dev = self.xyz.device
        ppll = D.PPLL(self.w, self.h, node_count, dev)
        ppll.clear()
        cap = None
        if capture:
            cap = dict(segment=torch.full((node_count,), -1, dtype=torch.int32, device=dev),
                       pixel=torch.empty(node_count, dtype=torch.int32, device=dev),
                       color=torch.empty((node_count, 4), dtype=torch.float32, device=dev))
        D.raster_strands(self.xyz, self.tan, self.th, self.idx, self.cam, self.light, self.sp, self.rp, self.vol,
                         self.res, self.smap, ppll, capture=cap)
        return ppll, cap


def _fragment_parity(s: Scene, ppll, cap, pairs=None, capacity=1 << 16):
    n = min(ppll.used(), ppll.node_count)
    seg = cap["segment"][:n].cpu().numpy().view(np.uint32)
    pix = cap["pixel"][:n].cpu().numpy().view(np.uint32)
    col = cap["color"][:n].cpu().numpy().astype(np.float64)
    nodes = ppll.nodes.cpu().numpy().view(np.uint32).reshape(-1, 3)[:n]
    ni = s.np_inputs()
    o = ORs.raster(ni["xyz"], ni["tan"], ni["th"], ni["idx"], s.cam, s.light, s.sp, s.rp, density=ni["dens"],
                   res=s.res, shadow_map=ni["smap"], pairs=pairs, capacity=capacity)
    kept = o["kept"]
    o_key = o["segment"][kept].astype(np.uint64) << np.uint64(32) | o["pixel"][kept].astype(np.uint64)
    # nodes a block reserved for fragments its shading discarded are never
    # linked nor captured (segment stays 0xffffffff)
    sel = seg != np.uint32(0xFFFFFFFF)
    if pairs is not None:
        sel &= np.isin(seg, np.asarray(pairs, np.uint32))
    g_key = seg[sel].astype(np.uint64) << np.uint64(32) | pix[sel].astype(np.uint64)
    assert len(np.unique(g_key)) == len(g_key), "a (segment, pixel) fragment inserted twice"
    common, gi, oi = np.intersect1d(g_key, o_key, return_indices=True)
    missing, extra = len(o_key) - len(common), len(g_key) - len(common)
    assert missing + extra <= max(2, 1e-4 * len(o_key)), (missing, extra, len(o_key))
    g_col = col[sel][gi]
    o_col = o["rgba"][kept][oi]
    err = float(np.abs(g_col - o_col).max()) if len(common) else 0.0
    assert err <= FRAG_TOL, err
    derr = float(np.abs(nodes[sel][gi, 1].view(np.float32).astype(np.float64) - o["depth"][kept][oi]).max())
    assert derr <= 1e-6, derr
    # the node colour is packUnorm4x8 of the captured colour
    assert np.array_equal(nodes[sel][:, 0], ORs.pack_unorm4x8(col[sel]))
    return dict(fragments=len(o_key), matched=len(common), missing=missing, extra=extra, max_err=err,
                max_depth_err=derr)


def _resolve_parity(ppll, background=None):
    out = ppll.resolve(background=background).cpu().numpy().astype(np.float64)
    heads = ppll.heads.cpu().numpy().view(np.uint32)
    nodes = ppll.nodes.cpu().numpy().view(np.uint32)
    used = min(ppll.used(), ppll.node_count)
    want = ORs.resolve(heads, nodes[:3 * max(used, 1)], ppll.width, ppll.height,
                       background=None if background is None else background.cpu().numpy())
    err = float(np.abs(out - want).max())
    assert err <= RESOLVE_TOL, err
    return err


@pytest.fixture(scope="module")
def c0(device):
    return Scene(synth.strands_10k(), (128, 128, 128), 512, 512, device, azimuth=30.0)


def test_shadow_map_matches_oracle(c0):
    ni = c0.np_inputs()
    want = ORs.shadow_depth(ni["xyz"], ni["idx"], _cm(c0.sc.light_vp), 1.0, (1024, 1024))
    got = ni["smap"]
    differ = got != want
    assert differ.mean() < 1e-4, differ.sum()
    assert (got < 1).sum() > 10000


def test_10k_strands_fragments_and_resolve(c0):
    ppll, cap = c0.raster(1 << 22)
    assert ppll.used() < ppll.node_count
    rep = _fragment_parity(c0, ppll, cap)
    print("10k strands raster", rep)
    assert rep["fragments"] > 100000
    # every node is reachable from exactly one head
    heads = ppll.heads.cpu().numpy().view(np.uint32)
    nodes = ppll.nodes.cpu().numpy().view(np.uint32).reshape(-1, 3)
    seen = np.zeros(ppll.used(), np.int32)
    for h0 in heads[heads != NULL][:2000]:
        h = h0
        while h != NULL:
            seen[h] += 1
            h = nodes[h, 2]
    assert seen.max() <= 1
    err = _resolve_parity(ppll)
    print("resolve max err", err)


def test_background_and_variants(c0, device):
    bg = torch.rand((c0.h, c0.w, 4), device=device)
    for rp in [params.render(renderer=0, shading_model=2), params.render(renderer=0, shading_model=3),
               params.render(renderer=3, lod_magnified_distance=0.0, lod_minified_distance=100.0)]:
        c0.rp = rp
        ppll, cap = c0.raster(1 << 22)
        rng = np.random.default_rng(2)
        pairs = rng.choice(c0.idx.shape[0] // 2, 3000, replace=False)
        _fragment_parity(c0, ppll, cap, pairs=pairs)
        _resolve_parity(ppll, background=bg)
    c0.rp = params.render(renderer=0)


def test_uneven_volume_resolution(device):
    """The LAO's per-axis cell strides and range tests at a non-cubic volume:
    fragments of a sample of segments against the oracle, and the resolve."""
    s = Scene(synth.strands_10k(), (96, 40, 150), 320, 240, device, azimuth=200.0)
    ppll, cap = s.raster(1 << 21)
    rng = np.random.default_rng(5)
    pairs = rng.choice(s.idx.shape[0] // 2, 4000, replace=False)
    rep = _fragment_parity(s, ppll, cap, pairs=pairs)
    print("uneven volume raster", rep)
    assert rep["fragments"] > 1000
    _resolve_parity(ppll)


def test_resolve_of_crafted_lists(device):
    """Node lists written on the host, 0-56 nodes per pixel (tails of up to
    40 past the 16 slots), with equal depths (the lowest slot blends first),
    depths of 0, -0, negative and NaN (never taken from the slots: slot 0's
    colour blends instead), and strong, distinct colours so that any change
    of blend order is far outside the tolerance."""
    rng = np.random.default_rng(11)
    w, h = 64, 32
    special = np.array([0.5, 0.5, 0.25, 0.0, -0.0, -0.3, np.nan, 0.75, 0.75, 1.0], np.float32)
    heads = np.full(w * h, NULL, np.uint32)
    nodes = []
    for p in range(w * h):
        for _ in range(int(rng.integers(0, 57))):
            if rng.random() < 0.6:
                d = special[rng.integers(len(special))]
            else:
                d = np.float32(rng.integers(1, 8) / 8)
            c = rng.integers(0, 256, 4)
            c[3] = rng.integers(96, 256)
            colour = int(c[0]) | int(c[1]) << 8 | int(c[2]) << 16 | int(c[3]) << 24
            nodes.append((colour, int(np.float32(d).view(np.uint32)), int(heads[p])))
            heads[p] = len(nodes) - 1
    nodes = np.asarray(nodes, np.uint32).reshape(-1)
    ppll = D.PPLL(w, h, len(nodes) // 3, device)
    ppll.heads.copy_(torch.from_numpy(heads.view(np.int32)))
    ppll.nodes.copy_(torch.from_numpy(nodes.view(np.int32)))
    out = ppll.resolve().cpu().numpy().astype(np.float64)
    want = ORs.resolve(heads, nodes, w, h)
    err = float(np.abs(out - want).max())
    assert err <= RESOLVE_TOL, err


def _crafted_ppll(device, w, h, lengths, rng, past_pool=()):
    """Host-written lists of the given lengths per pixel (random colours,
    depths in (0, 1] with ties); a pixel of past_pool links, midway, a node
    index one past the pool."""
    heads = np.full(w * h, NULL, np.uint32)
    parts, n_nodes = [], 0
    for p in range(w * h):
        n = int(lengths[p])
        if n == 0:
            continue
        ids = n_nodes + np.arange(n, dtype=np.uint32)
        colour = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32) | np.uint32(0x60000000)
        depth = (rng.integers(1, 64, n) / 64).astype(np.float32)
        nxt = np.concatenate([[NULL], ids[:-1]]).astype(np.uint32)   # the newest node (last id) is the head
        parts.append(np.stack([colour, depth.view(np.uint32), nxt], axis=1))
        heads[p] = ids[-1]
        n_nodes += n
    nodes = np.concatenate(parts).reshape(-1) if parts else np.zeros(3, np.uint32)
    count = max(n_nodes, 1)
    for p in past_pool:
        if heads[p] != NULL and lengths[p] > 4:
            nodes[3 * (int(heads[p]) - 3) + 2] = count   # the fourth node links one past the pool
    ppll = D.PPLL(w, h, count, device)
    ppll.heads.copy_(torch.from_numpy(heads.view(np.int32)))
    ppll.nodes.copy_(torch.from_numpy(nodes[:3 * count].view(np.int32)))
    return ppll, heads, nodes[:3 * count]


def test_resolve_long_lists_pool_ends_and_odd_sizes(device):
    """Lists at and past the resolve's 16 + 1024 node cap (the nodes beyond
    it are never read), lists that link a node index past the pool (they end
    there: the walk reads one node ahead, so this is its bound), and images
    that are not multiples of the 16x16 workgroup tile, against the oracle."""
    rng = np.random.default_rng(12)
    for w, h in ((17, 5), (1, 1), (33, 19)):
        lengths = rng.choice([0, 1, 3, 15, 16, 17, 40, 1039, 1040, 1041, 1100], size=w * h,
                             p=[0.2, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.05, 0.05, 0.05, 0.05])
        if w * h == 1:
            lengths[:] = 1100
        past = rng.choice(w * h, size=w * h // 8, replace=False)
        ppll, heads, nodes = _crafted_ppll(device, w, h, lengths, rng, past_pool=past)
        out = ppll.resolve().cpu().numpy().astype(np.float64)
        want = ORs.resolve(heads, nodes, w, h)
        err = float(np.abs(out - want).max())
        assert err <= RESOLVE_TOL, (w, h, err)


def test_pool_overflow_drops_without_out_of_bounds(c0, device):
    n = 5000
    guard = 64
    ppll = D.PPLL(c0.w, c0.h, n, device)
    ppll.nodes = torch.full((3 * (n + guard),), 0x5A5A5A5A, dtype=torch.int32, device=device)
    ppll.c.nodes = ppll.nodes.data_ptr()
    ppll.clear()
    D.raster_strands(c0.xyz, c0.tan, c0.th, c0.idx, c0.cam, c0.light, c0.sp, c0.rp, c0.vol, c0.res, c0.smap, ppll)
    assert ppll.used() > n
    nodes = ppll.nodes.cpu().numpy()
    assert (nodes[3 * n:] == 0x5A5A5A5A).all(), "write past the node pool"
    heads = ppll.heads.cpu().numpy().view(np.uint32)
    live = heads[heads != NULL]
    assert len(live) and (live < n).all()
    ppll.resolve()


def _list_nodes(ppll):
    """Node ids reachable from the heads (each list walked to its end)."""
    heads = ppll.heads.cpu().numpy().view(np.uint32)
    nxt = ppll.nodes.cpu().numpy().view(np.uint32).reshape(-1, 3)[:, 2]
    cur = heads[heads != NULL].astype(np.int64)
    seen = []
    while len(cur):
        assert (cur < ppll.node_count).all(), "a list links a node past the pool"
        seen.append(cur)
        cur = nxt[cur].astype(np.int64)
        cur = cur[cur != NULL]
    return np.concatenate(seen) if seen else np.zeros(0, np.int64)


def test_pool_between_linked_and_reserved_counts(c0, device):
    # The colour pass reserves a node id for every covered candidate, and
    # leaves the ids of fragments its shading discards unlinked, so the
    # counter runs ahead of the linked nodes.  With a pool between the two
    # counts, fragments whose id falls past its end are dropped though
    # enough ids were left unused below it.  Pinned here: the lists hold
    # exactly the captured nodes below the pool end, each once; every one of
    # them is a fragment of the unbounded run; the drops are no more than
    # the ids past the end; the resolve reads the lists as they are.
    big, cap_big = c0.raster(1 << 22)
    reserved = big.used()
    linked_big = int((cap_big["segment"][:reserved] != -1).sum())
    assert reserved > linked_big, "no discarded candidate at this view"
    key = lambda seg, pix: seg.astype(np.uint64) << np.uint64(32) | pix.astype(np.uint64)
    sel = cap_big["segment"][:reserved].cpu().numpy() != -1
    frags_big = key(cap_big["segment"][:reserved].cpu().numpy()[sel].view(np.uint32),
                    cap_big["pixel"][:reserved].cpu().numpy()[sel].view(np.uint32))
    n = linked_big + (reserved - linked_big) // 2
    small, cap = c0.raster(n)
    assert small.used() == reserved   # the same reservations, whatever the pool
    seg = cap["segment"].cpu().numpy()
    linked = np.flatnonzero(seg != -1)
    listed = _list_nodes(small)
    assert len(np.unique(listed)) == len(listed), "a node linked twice"
    assert np.array_equal(np.sort(listed), linked), "the lists and the captured nodes differ"
    frags = key(seg[linked].view(np.uint32), cap["pixel"].cpu().numpy()[linked].view(np.uint32))
    assert np.isin(frags, frags_big).all()
    assert 0 < len(linked) <= n and linked_big - len(linked) <= reserved - n
    _resolve_parity(small)


def test_size_limits_are_errors(c0, device):
    """Shadow maps beyond 16384 texels a side and volumes of 2^32 cells or
    more are refused before any launch (the taps use 32-bit offsets)."""
    ppll = D.PPLL(c0.w, c0.h, 1 << 16, device)
    wide = torch.zeros((1, 20000), dtype=torch.float32, device=device)
    with pytest.raises(capi.VkhrError) as e:
        D.raster_strands(c0.xyz, c0.tan, c0.th, c0.idx, c0.cam, c0.light, c0.sp, c0.rp, c0.vol, c0.res, wide, ppll)
    assert e.value.code == -1
    huge = dict(cells=c0.vol["cells"])
    with pytest.raises(capi.VkhrError) as e:
        D.raster_strands(c0.xyz, c0.tan, c0.th, c0.idx, c0.cam, c0.light, c0.sp, c0.rp, huge, (2048, 2048, 1100),
                         c0.smap, ppll)
    assert e.value.code == -3


def test_hybrid_raymarch_and_strands_share_the_ppll(c0, device):
    """Raymarched fragments go into the same list as the strands (the
    reference's subpass 0 strands + subpass 1 raymarch), then one resolve."""
    ppll, _ = c0.raster(1 << 22, capture=False)
    n_strand = ppll.used()
    rp = params.render(raycast_steps=256.0)
    D.raymarch_ppll(c0.vol, c0.res, c0.cam, c0.light, c0.sp, rp, ppll)
    rm = D.raymarch(c0.vol, c0.res, c0.cam, c0.light, c0.sp, rp, c0.w, c0.h, want_info=True)
    frags = (rm["info"][..., 0] == 1)
    assert ppll.used() - n_strand == int(frags.sum())
    # the raymarch nodes carry packUnorm4x8 of the raymarch colour
    nodes = ppll.nodes.cpu().numpy().view(np.uint32).reshape(-1, 3)[n_strand:ppll.used()]
    rgba = rm["rgba"].reshape(-1, 4).cpu().numpy()
    heads = ppll.heads.cpu().numpy().view(np.uint32)
    pix = np.nonzero(frags.reshape(-1).cpu().numpy())[0]
    head_nodes = heads[pix]
    is_rm = head_nodes >= n_strand                     # raymarch node inserted last: it is the head
    assert is_rm.all()
    assert np.array_equal(nodes[head_nodes - n_strand, 0], ORs.pack_unorm4x8(rgba[pix]))
    _resolve_parity(ppll)


def _lod(rp, cam):
    """lod() of scheme.glsl:7-11 for Hybrid_LoD, in float64 from the float32
    parameters (as the kernels and the oracles take it)."""
    e0, e1 = float(np.float32(rp.lod_magnified_distance)), float(np.float32(rp.lod_minified_distance))
    t = min(max((float(np.float32(cam.look_at_distance)) - e0) / (e1 - e0), 0.0), 1.0)
    return t * t * (3.0 - 2.0 * t)


def _capture_table(ppll, cap):
    n = min(ppll.used(), ppll.node_count)
    seg = cap["segment"][:n].cpu().numpy().view(np.uint32)
    pix = cap["pixel"][:n].cpu().numpy().view(np.uint32)
    col = cap["color"][:n].cpu().numpy().astype(np.float64)
    sel = seg != np.uint32(NULL)
    key = seg[sel].astype(np.uint64) << np.uint64(32) | pix[sel].astype(np.uint64)
    order = np.argsort(key)
    return key[order], col[sel][order]


def test_hybrid_lod_scales_each_pass(c0, device):
    """Renderer Hybrid_LoD at a level between the LoD distances: the strand
    coverage is scaled by 1 - lod after the discard test (strand.frag:39-42),
    so the fragment set is the rasterizer's; the raymarch coverage by lod
    (volume.frag:48); colour and depth are unchanged.  The oracle is pinned
    by the same ratios against its pure-rasterizer and pure-raymarcher runs
    (lod = 0.156 here, so a swapped lod / 1 - lod fails), the GPU against the
    oracle at the hybrid level, and both passes then share one PPLL."""
    d = float(c0.cam.look_at_distance)
    rph = params.render(renderer=3, raycast_steps=256.0, lod_magnified_distance=0.75 * d,
                        lod_minified_distance=1.75 * d)
    lod = _lod(rph, c0.cam)
    assert 0.1 < lod < 0.2, lod
    ni = c0.np_inputs()
    # the oracle's own scaling, on sampled segments
    rng = np.random.default_rng(11)
    pairs = rng.choice(c0.idx.shape[0] // 2, 1500, replace=False)
    kw = dict(density=ni["dens"], res=c0.res, shadow_map=ni["smap"], pairs=pairs)
    o0 = ORs.raster(ni["xyz"], ni["tan"], ni["th"], ni["idx"], c0.cam, c0.light, c0.sp, params.render(renderer=0), **kw)
    oh = ORs.raster(ni["xyz"], ni["tan"], ni["th"], ni["idx"], c0.cam, c0.light, c0.sp, rph, **kw)
    assert np.array_equal(o0["kept"], oh["kept"]) and np.array_equal(o0["pixel"], oh["pixel"])
    k = o0["kept"]
    assert k.sum() > 1000
    assert np.array_equal(oh["rgba"][k, :3], o0["rgba"][k, :3])
    k &= o0["rgba"][:, 3] < 1.0                      # alpha below the clamp
    assert np.abs(oh["rgba"][k, 3] - o0["rgba"][k, 3] * (1.0 - lod)).max() <= 1e-12
    # the GPU at the hybrid level, and against its own rasterizer run
    saved = c0.rp
    try:
        c0.rp = params.render(renderer=0)
        p0, cap0 = c0.raster(1 << 22)
        c0.rp = rph
        ph, caph = c0.raster(1 << 22)
        rep = _fragment_parity(c0, ph, caph)
        print("hybrid strands", rep)
    finally:
        c0.rp = saved
    k0, col0 = _capture_table(p0, cap0)
    kh, colh = _capture_table(ph, caph)
    assert np.array_equal(k0, kh)
    assert np.abs(colh[:, :3] - col0[:, :3]).max() <= 1e-6
    under = col0[:, 3] < 1.0
    assert np.abs(colh[under, 3] - col0[under, 3] * (1.0 - lod)).max() <= 1e-6

    # raymarch: the oracle's ratio, the GPU against the oracle, the GPU ratio
    rp2 = params.render(renderer=2, raycast_steps=256.0)
    pix = rng.choice(c0.w * c0.h, 4000, replace=False)
    r2 = OR.raymarch(ni["dens"], c0.vol["tangents"].cpu().numpy(), c0.res, c0.cam, c0.light, c0.sp, rp2, c0.w,
                     c0.h, pixels=pix)
    rh = OR.raymarch(ni["dens"], c0.vol["tangents"].cpu().numpy(), c0.res, c0.cam, c0.light, c0.sp, rph, c0.w,
                     c0.h, pixels=pix)
    f = r2["info"][:, 0] == 1
    assert f.sum() > 100 and np.array_equal(f, rh["info"][:, 0] == 1)
    unclamped = f & (r2["rgba"][:, 3] < 1.0)
    assert np.array_equal(rh["rgba"][f, :3], r2["rgba"][f, :3])
    assert np.abs(rh["rgba"][unclamped, 3] - r2["rgba"][unclamped, 3] * lod).max() <= 1e-12

    def render(rp):
        out = D.raymarch(c0.vol, c0.res, c0.cam, c0.light, c0.sp, rp, c0.w, c0.h, want_depth=True, want_info=True)
        return {key: v.cpu().numpy() for key, v in out.items() if v is not None}

    g2, gh = render(rp2), render(rph)
    print("hybrid raymarch", compare_raymarch(gh, dict(density=ni["dens"], tangents=c0.vol["tangents"].cpu().numpy()),
                                              c0.res, c0.cam, c0.light, c0.sp, rph, c0.w, c0.h))
    gf = g2["info"][..., 0] == 1
    assert np.array_equal(gf, gh["info"][..., 0] == 1)
    a2, ah = g2["rgba"][gf].astype(np.float64), gh["rgba"][gf].astype(np.float64)
    assert np.array_equal(ah[:, :3], a2[:, :3]) and np.array_equal(gh["depth"][gf], g2["depth"][gf])
    sel = a2[:, 3] < 1.0
    assert np.abs(ah[sel, 3] - a2[sel, 3] * lod).max() <= 1e-6

    # both passes into one list, one resolve
    n_strand = ph.used()
    D.raymarch_ppll(c0.vol, c0.res, c0.cam, c0.light, c0.sp, rph, ph)
    assert ph.used() - n_strand == int(gf.sum())
    print("hybrid resolve", _resolve_parity(ph))


def test_136k_strands_full_scale(device):
    """136,320 strands (1,635,840 segments) at 1920x1080, 1024^2 shadow map,
    k=16 and a 32M-node pool: the oracle checks every fragment of every
    segment (16.7 M) and the whole resolve."""
    s = Scene(synth.strands_136k(), (256, 256, 256), 1920, 1080, device)
    ppll, cap = s.raster(1 << 25)
    assert ppll.used() < ppll.node_count
    rep = _fragment_parity(s, ppll, cap, capacity=ppll.used())
    assert rep["fragments"] > 16_000_000
    print("136k strands raster", rep, "nodes", ppll.used())
    err = _resolve_parity(ppll)
    print("136k strands resolve", err)
